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
