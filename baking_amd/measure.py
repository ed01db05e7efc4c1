"""Acceptance aggregates of a simulation, reduced on the GPU (``bk_measure``).

The reference's run measurements (experiments/recording.py:46-112) and the
statistics of ``MpmSimulator.CheckState`` (physics/mpm.py:292-308): particle
volume sum V0 det F, capacity-weighted mean temperature, water mass, water
lost, heat exchanged, thermal energy sum capacity * T, gas content sum V0 N and
total mass, as f64 sums per dough; for a ``BatchSimulator`` one entry per dough.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_SUMS = ("particle_volume_m3", "thermal_energy_j", "heat_capacity_j_per_k", "water_mass_kg", "water_lost_kg",
         "heat_exchange_j", "gas_total_mol", "total_mass")


def _batch_args(sim):
    if hasattr(sim, "doughs"):  # BatchSimulator
        return sim._scene, sim._state_struct(sim._cur), sim.particle_count, sim.dough_count
    return sim._scene, sim._state_struct(sim.current_state), sim.particle_data.particle_count, 1


def measure(sim):
    """{name: float64 array of one value per dough} (a single simulator gives arrays of length 1)."""
    import torch
    scene, state, n, nd = _batch_args(sim)
    out = torch.empty(nd * len(_SUMS), dtype=torch.float64, device="cuda")
    _native.check(_native.load().bk_measure(ctypes.byref(scene), ctypes.byref(state), n, nd, _native.ptr(out),
                                            _native.stream_ptr()), "bk_measure")
    sums = out.view(nd, len(_SUMS)).cpu().numpy()
    res = {name: sums[:, k].copy() for k, name in enumerate(_SUMS)}
    cap = res["heat_capacity_j_per_k"]
    res["temperature_mean_k"] = np.divide(res["thermal_energy_j"], cap, out=np.zeros_like(cap), where=cap > 0)
    return res


def _static(sim):
    """Per simulator, once: the surface particles (face owners) and the core particles of the reference recorder
    (experiments/recording.py:80-85: initial sample positions within 0.4 of the largest distance from the mesh
    centre, or the closest ones), and the initial water and energy."""
    import torch
    cached = getattr(sim, "_measure_static", None)
    if cached is not None:
        return cached
    dev = sim.current_state.particle_positions.device
    out = {}
    if sim.transport is not None:
        n = sim.particle_data.particle_count
        surface = np.zeros(n, dtype=bool)
        surface[sim.transport.owners.cpu().numpy()] = True
        pos = np.asarray(sim.particle_samples.particle_positions)
        radius = np.linalg.norm(pos - np.mean(sim.tet_mesh.vertex_positions, axis=0), axis=1)
        core = radius < 0.4 * radius.max()
        if not core.any():
            core = radius == radius.min()
        t = sim.transport.config
        V0 = np.asarray(sim.particle_samples.reference_volumes, dtype=np.float64)
        out.update(surface=torch.as_tensor(np.flatnonzero(surface), device=dev),
                   core=torch.as_tensor(np.flatnonzero(core), device=dev),
                   initial_water=float(t.water_density * t.initial_saturation * V0.sum()),
                   initial_energy=float(t.initial_temperature_k) * float(sim.transport.capacities.double().sum()))
    out["V0"] = torch.as_tensor(np.asarray(sim.particle_samples.reference_volumes, dtype=np.float64), device=dev)
    out["tets"] = torch.as_tensor(np.asarray(sim.tet_mesh.tet_indices, dtype=np.int64), device=dev)
    sim._measure_static = out
    return out


def Measure(sim):
    """The reference recorder's measurement of a simulator's current state (experiments/recording.py:46-112), reduced
    on the GPU: the sums of ``bk_measure``, the extremes and means as device reductions, the mesh after UpdateMesh,
    all copied to the host in one transfer.  Same keys and definitions as the reference's dict."""
    import torch
    sums = measure(sim)
    st = _static(sim)
    sim.UpdateMesh()
    s = sim.current_state
    verts = sim.vertex_positions.double()
    F = s.deformation_gradients.double()
    if not bool(torch.isfinite(verts).all()) or not bool(torch.isfinite(F).all()):
        raise ValueError("Cannot record nonfinite simulation state")
    corners = verts[st["tets"]]
    signed = torch.linalg.det(corners[:, 1:] - corners[:, :1]) / 6.0
    J = torch.linalg.det(F)
    names, vals = [], []

    def put(name, value):
        names.append(name)
        vals.append(value.double().reshape(()))

    m = s.material_state
    ding = m is not None and hasattr(m, "porosities")
    if ding:
        por = m.porosities
        put("temperature_k", m.temperatures.mean())
        put("porosity_min", por.min())
        put("porosity_mean", por.double().mean())
        put("porosity_max", por.max())
        put("yield_stress_pa", m.yield_stresses[0])
        put("viscosity_pa3_s", m.viscosities[0])
        put("gelled_fraction", m.gelled.double().mean())
        if sim.material.ding_config.gas_enabled:
            pr = m.gas_pressures
            put("gas_pressure_min_pa", pr.min())
            put("gas_pressure_max_pa", pr.max())
            put("gas_pressure_mean_pa", pr.mean())
            put("pore_volume_m3", (st["V0"] * J * por.double()).sum())
    t = s.transport_state
    if t is not None:
        T, sat = t.temperatures, t.saturations
        put("temperature_min_k", T.min())
        put("temperature_max_k", T.max())
        put("temperature_surface_k", T[st["surface"]].mean())
        put("temperature_core_k", T[st["core"]].mean())
        put("saturation_min", sat.min())
        put("saturation_max", sat.max())
        put("saturation_mean", sat.mean())
    lo, hi = verts.min(dim=0).values, verts.max(dim=0).values
    put("length_x_m", hi[0] - lo[0])
    put("width_z_m", hi[2] - lo[2])
    put("height_y_m", hi[1] - lo[1])
    put("mesh_volume_m3", signed.abs().sum())
    put("inverted_tets", (signed <= 0).sum())
    put("invalid_particles", (J <= 0).sum())
    host = dict(zip(names, torch.stack(vals).cpu().tolist()))
    out = {k: host[k] for k in names if k not in ("length_x_m", "width_z_m", "height_y_m", "mesh_volume_m3", "inverted_tets",
                                                     "invalid_particles")}
    if ding and sim.material.ding_config.gas_enabled:
        out["gas_total_mol"] = float(sums["gas_total_mol"][0])
    if t is not None:
        losses, energy, exchange = (float(sums[k][0]) for k in ("water_lost_kg", "thermal_energy_j", "heat_exchange_j"))
        water = float(sums["water_mass_kg"][0])
        out.update(temperature_mean_k=float(sums["temperature_mean_k"][0]), water_mass_kg=water, water_lost_kg=losses,
                   water_balance_error_kg=water + losses - st["initial_water"], thermal_energy_j=energy,
                   heat_exchange_j=exchange, energy_balance_error_j=energy - st["initial_energy"] - exchange)
    out.update(baking_time_s=sim.baking_time, phase="settling" if sim.IsSettling else "baking", step=sim.step_count,
               time_s=sim.step_count * sim.simulation_config.time_step, length_x_m=host["length_x_m"],
               width_z_m=host["width_z_m"], height_y_m=host["height_y_m"], particle_volume_m3=float(sums["particle_volume_m3"][0]),
               mesh_volume_m3=host["mesh_volume_m3"], inverted_tets=int(host["inverted_tets"]),
               invalid_particles=int(host["invalid_particles"]))
    return out
