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
