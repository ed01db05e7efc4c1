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


def oracle_from_spec(spec, precision="f32"):
    sim = O.OracleSimulator(spec.tet_mesh, spec.samples, spec.simulation, spec.material, spec.initial_velocity, spec.ding,
                            spec.transport, precision=precision)
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


def field_scales(ref, dx):
    """Magnitude floors for fields whose exact value can be ~0 while its terms are not.

    C = (4/dx^2) sum_i w_i v_i (x_i - x_p)^T: its terms scale like 4 max|v| / dx, so a
    rigidly moving body (C ~ 0) is compared at that scale.
    """
    vmax = float(np.abs(np.asarray(ref["v"], dtype=np.float64)).max(initial=0.0))
    return {"C": 4.0 * vmax / dx}


def field_errors(a, b, names=None, floors=None):
    """Scale-relative max error per field: max|a - b| / max(max|b|, floor, tiny)."""
    errs = {}
    floors = floors or {}
    for k in (names or sorted(set(a) & set(b))):
        x = np.asarray(a[k], dtype=np.float64).reshape(-1)
        y = np.asarray(b[k], dtype=np.float64).reshape(-1)
        scale = max(np.abs(y).max(initial=0.0), floors.get(k, 0.0), 1e-30)
        errs[k] = float(np.abs(x - y).max(initial=0.0) / scale)
    return errs


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


def three_way(spec, steps, rtol=1e-5, factor=4.0, check_every=1, log=print, solver=None):
    """Run the GPU path, the FP32 oracle and the FP64 oracle side by side.

    After each checked substep every field must satisfy
        err(gpu, oracle32) <= max(rtol, factor * err(oracle32, oracle64)),
    i.e. the GPU agrees with the reference algorithm to 1e-5 (scale-relative),
    or, where the field is rounding-dominated (stress of a nearly undeformed
    body), to within a few times the reference's own FP32 rounding error.
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
        r32, r64 = oracle_fields(o32), oracle_fields(o64)
        floors = field_scales(r32, o32.dx)
        e_gpu = field_errors(gpu_fields(gpu), r32, floors=floors)
        e_ref = field_errors(r32, r64, floors=floors)
        bad = {k: (e_gpu[k], e_ref[k]) for k in e_gpu if e_gpu[k] > max(rtol, factor * e_ref[k])}
        if log:
            log(f"{spec.name} step {step}: worst gpu/o32 {max(e_gpu.values()):.2e}; "
                + ", ".join(f"{k} {e_gpu[k]:.1e}/{e_ref[k]:.1e}" for k in sorted(e_gpu) if e_gpu[k] > 1e-7))
        assert not bad, f"{spec.name} step {step}: fields outside tolerance (gpu/o32, o32/o64): {bad}"
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
    """Copy the GPU simulator's state, dense grid masses, flags and step count into the oracle ``o``."""
    s = sim.current_state
    host = lambda t: t.detach().cpu().numpy().copy()  # noqa: E731
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
