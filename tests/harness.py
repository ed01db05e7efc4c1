"""Shared test helpers: build the oracle and the GPU simulator from one scene and compare fields.

The oracle (CPU restatement of the reference) is the checker; the GPU path is
``baking_amd``'s fused HIP substep.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402


def oracle_from_spec(spec, precision="f32", library=None, sticky_floor=False):
    """The oracle simulator of a scene (``library``: a timing build's path; default the checker build)."""
    sim = O.OracleSimulator(spec.tet_mesh, spec.samples, spec.simulation, spec.material, spec.initial_velocity, spec.ding,
                            spec.transport, precision=precision, library=library, sticky_floor=sticky_floor)
    st = sim.state
    if spec.water_masses is not None:
        st.transport["W"][...] = spec.water_masses
    if spec.porosities is not None:
        st.ding["porosity"][...] = spec.porosities
    if spec.gas_moles is not None:
        st.ding["moles"][...] = spec.gas_moles
    if spec.velocities is not None:
        st.v[...] = spec.velocities
    if spec.affine_velocities is not None:
        st.C[...] = spec.affine_velocities
    if spec.temperatures is not None:
        st.transport["T"][...] = spec.temperatures
    return sim


def gpu_from_spec(spec):
    from baking_amd.scenes import build_simulator
    return build_simulator(spec)


def gpu_fields(sim):
    s = sim.current_state
    out = dict(x=s.particle_positions, v=s.particle_velocities, F=s.deformation_gradients, C=s.affine_velocities)
    m = s.material_state
    if m is not None:
        out.update(Fe=m.elastic_gradients, Fp=m.plastic_gradients, tau=m.kirchhoff_stresses, porosity=m.porosities,
                   T_mat=m.temperatures, gelled=m.gelled, yield_stress=m.yield_stresses, viscosity=m.viscosities,
                   residual=m.return_residuals, moles=m.gas_moles, pressure=m.gas_pressures, density=m.reference_densities,
                   porosity_residual=m.porosity_residuals)
    t = s.transport_state
    if t is not None:
        out.update(T=t.temperatures, W=t.water_masses, s=t.saturations, losses=t.water_losses, exchanges=t.heat_exchanges,
                   masses=t.particle_masses)
    out["node_m"] = sim.grid_state.node_masses
    out["node_v"] = sim.grid_state.node_velocities
    return {k: v.detach().cpu().numpy() for k, v in out.items()}


def oracle_fields(sim):
    s = sim.state
    out = dict(x=s.x, v=s.v, F=s.F, C=s.C)
    if s.ding is not None:
        d = s.ding
        out.update(Fe=d["Fe"], Fp=d["Fp"], tau=d["tau"], porosity=d["porosity"], T_mat=d["T"], gelled=d["gelled"],
                   yield_stress=d["yield_stress"], viscosity=d["viscosity"], residual=d["residual"], moles=d["moles"],
                   pressure=d["pressure"], density=d["density"], porosity_residual=d["porosity_residual"])
    if s.transport is not None:
        t = s.transport
        out.update(T=t["T"], W=t["W"], s=t["s"], losses=t["losses"], exchanges=t["exchanges"], masses=t["masses"])
    out["node_m"] = sim.node_m
    out["node_v"] = sim.node_v
    return out


def field_scales(ref, orc):
    """Physical magnitude floors: each field is compared at max(max|field|, floor), never at less.

    ``ref`` holds the fields of the oracle simulator ``orc`` (its configs and step count set the floors).
    A field whose exact value is ~0 while its terms are not (a freely falling body's stress, a rigid body's
    affine velocity) would otherwise be compared at the size of its own rounding noise.  Floors:

    * v, node_v: |g| t, the speed gravity gives in the elapsed time t, and tau_floor / (rho c), the particle
      speed of an elastic wave carrying the stress floor below (c = sqrt((lambda + 2 mu) / rho)): a velocity
      error well below it implies a strain-rate, and so a stress, error well below the stress floor.  node_v
      errors are weighted by the node's mass over the largest node mass (``field_errors``' weights): p/m on a
      fringe node that holds 1e-9 of a particle's mass divides rounding noise by that mass, and G2P gives such
      a node a proportionally small weight, so it is its momentum error, not its velocity error, that reaches
      the particles;
    * C = (4/dx^2) sum_i w_i v_i (x_i - x_p)^T: its terms scale like 4 |v| / dx;
    * tau: the yield stress of the Ding dough (the stress at which it flows), or 1e-3 mu for neo-Hookean
      (the stress of a strain of 1e-3), and the weight rho |g| L of a column of height L = 2 dx;
    * pressure: the ambient pore pressure; moles: their largest value;
    * residual: 1, the unit term of the return map's equation gamma + 1 - q/(1+d gamma)^2 - p/(1+m gamma)^2;
    * porosity_residual, porosity: the initial porosity n0 (both are in porosity units);
    * losses, exchanges: 1e-6 of the water the dough holds and of its heat content C T0 (both start at 0).
    """
    sim, mat = orc.sim, orc.mat
    t = max(orc.step_count, 1) * float(sim.time_step)
    g = float(np.linalg.norm(np.asarray(sim.gravity, dtype=np.float64)))
    stress = float(orc.ding_cfg.yield_stress) if orc.ding_cfg is not None else 1e-3 * float(mat.shear_modulus)
    stress = max(stress, float(mat.density) * g * 2.0 * float(sim.grid_spacing))
    acoustic = stress / (float(mat.density) * float(mat.wave_speed))
    vmax = float(np.abs(np.asarray(ref["v"], dtype=np.float64)).max(initial=0.0))
    vscale = max(vmax, g * t, acoustic)
    floors = {"v": vscale, "node_v": vscale, "C": 4.0 * vscale / float(sim.grid_spacing), "tau": stress}
    if orc.ding_cfg is not None:
        c = orc.ding_cfg
        floors["pressure"] = float(c.gas_ambient_pressure_pa)
        floors["porosity"] = floors["porosity_residual"] = max(float(c.porosity), 1e-3)
        floors["residual"] = 1.0
        floors["moles"] = float(np.abs(np.asarray(ref["moles"], dtype=np.float64)).max(initial=0.0))
    if orc.tr is not None:
        floors["losses"] = 1e-6 * float(np.abs(np.asarray(ref["W"], dtype=np.float64)).sum())
        floors["exchanges"] = 1e-6 * float(np.sum(orc.capacities)) * float(orc.tr.initial_temperature_k)
    return floors


def field_errors(a, b, names=None, floors=None, weights=None):
    """Scale-relative max error per field: max_i w_i |a_i - b_i| / max(max|b|, floor, tiny), w_i = 1 unless
    ``weights`` gives a per-row weight for the field (node_v: node mass / largest node mass of ``b``)."""
    errs = {}
    floors = floors or {}
    if weights is None and "node_v" in b and "node_m" in b:
        m = np.asarray(b["node_m"], dtype=np.float64).reshape(-1)
        weights = {"node_v": m / max(m.max(initial=0.0), 1e-300)}
    weights = weights or {}
    for k in (names or sorted(set(a) & set(b))):
        x = np.asarray(a[k], dtype=np.float64)
        y = np.asarray(b[k], dtype=np.float64)
        scale = max(np.abs(y).max(initial=0.0), floors.get(k, 0.0), 1e-30)
        diff = np.abs(x - y).reshape(len(y), -1) if y.ndim else np.abs(x - y).reshape(1, 1)
        if k in weights:
            diff = diff * weights[k].reshape(-1, 1)
        errs[k] = float(diff.max(initial=0.0) / scale)
    return errs


ROUNDING_CAP = 1e-2


def rate_fields(e_gpu, e_ref, rtol=1e-5, factor=4.0, cap=ROUNDING_CAP):
    """Per field: 'rtol' when err(gpu, o32) <= rtol; 'fp32' when it is within factor x err(o32, o64) (the
    reference's own FP32 rounding) and that allowance is at most ``cap``; otherwise 'FAIL'.  So a field passes
    either at 1e-5 of its physical scale or at a few FP32 roundings, and never at 1e-2 relative error or more.
    (The FP32 rounding of the reference algorithm itself reaches ~1e-3 of the physical scale in the stiff
    small sphere: 5 ulp of strain at bulk modulus 1.4 MPa is ~1 Pa of stress against the 811 Pa yield
    stress, and its acoustic waves carry that into the velocities.)"""
    out = {}
    for k, e in e_gpu.items():
        if e <= rtol:
            out[k] = "rtol"
        elif e <= min(factor * e_ref[k], cap):
            out[k] = "fp32"
        else:
            out[k] = "FAIL"
    return out


def check_fields(tag, got, o32, o64, log=print, rtol=1e-5, factor=4.0, cap=ROUNDING_CAP):
    """Fields ``got`` against the FP32 oracle simulator ``o32`` under ``rate_fields``, with the FP64 oracle
    ``o64`` at the same substep giving the FP32 rounding allowance.  Logs each field's error, the oracle's own
    FP32/FP64 gap and the branch it passed by; returns (errors, gaps)."""
    r32, r64 = oracle_fields(o32), oracle_fields(o64)
    floors = field_scales(r32, o32)
    e_gpu = field_errors(got, r32, floors=floors)
    e_ref = field_errors(r32, r64, floors=floors)
    outcome = rate_fields(e_gpu, e_ref, rtol, factor, cap)
    if log:
        log(f"{tag}: worst {max(e_gpu.values()):.2e}; "
            + ", ".join(f"{k} {e_gpu[k]:.1e}/{e_ref[k]:.1e} {outcome[k]}" for k in sorted(e_gpu) if e_gpu[k] > 1e-7))
    bad = {k: (e_gpu[k], e_ref[k]) for k, v in outcome.items() if v == "FAIL"}
    assert not bad, f"{tag}: fields outside tolerance (gpu/o32, o32/o64): {bad}"
    return e_gpu, e_ref


def check_same(tag, a, b, o32, o64, rtol=1e-5, factor=4.0, cap=ROUNDING_CAP):
    """Two GPU runs of one scene (batch vs alone, tuned vs untuned, guided vs exact) agree under the same rule,
    with the FP32/FP64 oracle gap of that scene at that substep as the rounding allowance."""
    r32, r64 = oracle_fields(o32), oracle_fields(o64)
    floors = field_scales(r32, o32)
    e = field_errors(a, b, floors=floors)
    e_ref = field_errors(r32, r64, floors=floors)
    outcome = rate_fields(e, e_ref, rtol, factor, cap)
    bad = {k: (e[k], e_ref[k]) for k, v in outcome.items() if v == "FAIL"}
    assert not bad, f"{tag}: runs differ beyond tolerance: {bad}"


class OraclePoint:
    """One-particle Ding state advanced by the oracle's UpdateDing (+ RefreshGas when gas is on).

    ``step(F_total, dt)`` mirrors the reference's single-point driver
    (tests/test_ding.py:23-29): the new total gradient is F_total and
    C = (F_total F_old^-1 - I) / dt.
    """

    def __init__(self, material, ding, initial_temperature_k=293.15, precision="f32"):
        self.k = O.Kernels(precision)
        R = self.k.R
        self.R, self.mat, self.ding = R, material, ding
        T0 = initial_temperature_k
        gel = int(ding.temperature_enabled and T0 >= ding.temperature_knots_k[3])
        yf = float(np.interp(T0, ding.temperature_knots_k, ding.yield_factors)) if ding.temperature_enabled else 1.0
        vf = float(np.interp(T0, ding.temperature_knots_k, ding.viscosity_factors)) if ding.temperature_enabled else 1.0
        self.visc = np.array([ding.viscosity], R)
        self.F = np.eye(3, dtype=R)[None].copy()
        self.s = dict(Fe=np.eye(3, dtype=R)[None].copy(), Fp=np.eye(3, dtype=R)[None].copy(), tau=np.zeros((1, 3, 3), R),
                      residual=np.zeros(1, R), porosity=np.full(1, ding.porosity, R), T=np.full(1, T0, R),
                      gelled=np.full(1, gel, np.int32), yield_stress=np.full(1, ding.yield_stress * yf, R),
                      viscosity=np.full(1, float(R(ding.viscosity)) * vf, R), porosity_residual=np.zeros(1, R),
                      moles=np.full(1, ding.gas_initial_pressure_pa * ding.porosity / (O.GAS_CONSTANT * T0)
                                    if ding.gas_enabled else 0.0),
                      pressure=np.full(1, ding.gas_initial_pressure_pa if ding.gas_enabled else 0.0))
        self.flags = np.zeros(3, np.int32)

    def step(self, F_total, dt=0.01):
        R = self.R
        inc = np.asarray(F_total, np.float64) @ np.linalg.inv(self.F[0].astype(np.float64))
        C = ((inc - np.eye(3)) / dt).astype(R)[None].copy()
        Fn = np.asarray(F_total).astype(R)[None].copy()
        o, d = self.s, self.ding
        n = {k: v.copy() for k, v in o.items()}
        self.k.UpdateDing(1, o["Fe"], o["porosity"], o["gelled"], Fn, C, n["Fe"], n["Fp"], n["tau"], n["residual"], n["porosity"],
                          n["T"], n["gelled"], n["yield_stress"], n["viscosity"], self.visc, self.flags, n["porosity_residual"],
                          int(d.gas_enabled), self.mat.shear_modulus, self.mat.lame_lambda, d.yield_stress, d.porosity,
                          int(d.temperature_enabled), np.asarray(d.temperature_knots_k, R), np.asarray(d.yield_factors, R),
                          np.asarray(d.viscosity_factors, R), dt, o["T"])
        if d.gas_enabled:
            self.k.RefreshGas(1, o["moles"], n["moles"], Fn, n["porosity"], n["pressure"], n["T"])
        self.s, self.F = n, Fn
        return n


def three_way(spec, steps, rtol=1e-5, factor=4.0, check_every=1, log=print, solver=None, cap=ROUNDING_CAP):
    """Run the GPU path, the FP32 oracle and the FP64 oracle side by side.

    After each checked substep every field must satisfy (``rate_fields``)
        err(gpu, oracle32) <= rtol                                   ("rtol"), or
        err(gpu, oracle32) <= min(factor * err(oracle32, oracle64), cap)   ("fp32"),
    with errors scale-relative to max(max|field|, physical floor) (``field_scales``): the GPU agrees with
    the reference algorithm to 1e-5 of the field's physical scale, or, where FP32 rounding of the
    reference itself is larger, to within a few of those roundings and never worse than ``cap``.
    Returns the last errors (gpu vs oracle32, oracle32 vs oracle64).
    """
    gpu = gpu_from_spec(spec)
    if solver is not None:
        gpu.SetReturnSolver(solver)
    o32 = oracle_from_spec(spec, "f32")
    o64 = oracle_from_spec(spec, "f64")
    e_gpu = e_ref = None
    for step in range(1, steps + 1):
        gpu.Step()
        o32.Step()
        o64.Step()
        if step % check_every and step != steps:
            continue
        e_gpu, e_ref = check_fields(f"{spec.name} step {step}", gpu_fields(gpu), o32, o64, log=log, rtol=rtol,
                                    factor=factor, cap=cap)
    return e_gpu, e_ref


_DING_FIELDS = (("Fe", "elastic_gradients"), ("Fp", "plastic_gradients"), ("tau", "kirchhoff_stresses"),
                ("residual", "return_residuals"), ("porosity", "porosities"), ("T", "temperatures"), ("gelled", "gelled"),
                ("yield_stress", "yield_stresses"), ("viscosity", "viscosities"), ("moles", "gas_moles"),
                ("pressure", "gas_pressures"), ("porosity_residual", "porosity_residuals"), ("density", "reference_densities"))
_TRANSPORT_FIELDS = (("T", "temperatures"), ("W", "water_masses"), ("s", "saturations"), ("losses", "water_losses"),
                     ("exchanges", "heat_exchanges"), ("masses", "particle_masses"))


def oracle_on_gpu_state(spec, sim):
    """An FP32 oracle holding the GPU simulator's current state (copied to the host), dense grid masses, flags
    and step count: the reference's host-side computations (CheckState) on exactly the GPU's state."""
    return load_gpu_state(oracle_from_spec(spec, "f32"), sim)


def load_gpu_state(o, sim):
    """Copy the GPU simulator's state, dense grid masses, flags and step count into the oracle ``o``.
    float32 fields become the oracle's working precision (float64 for the FP64 oracle); float64 and int
    fields are kept as they are."""
    s = sim.current_state
    R = o.R

    def host(t):
        a = t.detach().cpu().numpy()
        return np.ascontiguousarray(a.astype(R) if a.dtype == np.float32 else a.copy())

    o.state.x, o.state.v = host(s.particle_positions), host(s.particle_velocities)
    o.state.F, o.state.C = host(s.deformation_gradients), host(s.affine_velocities)
    if s.material_state is not None:
        for key, name in _DING_FIELDS:
            o.state.ding[key] = host(getattr(s.material_state, name))
    if s.transport_state is not None:
        for key, name in _TRANSPORT_FIELDS:
            o.state.transport[key] = host(getattr(s.transport_state, name))
    o.node_m = host(sim.grid_state.node_masses).reshape(-1)
    flags = sim._read_flags()
    o.flags[:] = flags[:2]
    o.ding_flags[:] = flags[2:5]
    o.step_count = sim.step_count
    return o
