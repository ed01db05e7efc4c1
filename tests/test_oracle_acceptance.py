"""The acceptance aggregates of the oracle (reference experiments/recording.py:46-112) and the committed
1000-substep oracle fixture of the small sphere scene.

* ``OracleSimulator.Measure`` against closed forms on an insulated, uniformly heated dough cube translating
  freely: total particle volume sum V0 det F = sum V0 (rigid motion keeps J = 1), capacity-weighted mean
  temperature = the uniform temperature, water = rho_w s0 sum V0, gas = sum V0 p0 n0 / (R T0), no water
  lost and no heat exchanged; then a planted per-particle temperature field, whose capacity-weighted mean
  is computed independently.
* ``tests/golden/sphere_small_1000.npz`` (written by ``make_sphere_small_1000.py``): its first 100 substeps
  are re-derived here by both oracles, so the fixture the GPU acceptance test reads is the oracle's run.
"""
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from tests.harness import O, oracle_from_spec

GOLDEN = Path(__file__).resolve().parent / "golden"


def _insulated_translating_cube(T1=330.0):
    from baking_amd import scenes
    spec = scenes.demo_transport(heat_transfer=0.0, gas=True)
    spec.transport = replace(spec.transport, initial_temperature_k=T1, conductivity=0.0, tray_heat_transfer=0.0,
                             surface_water_loss=0.0, water_diffusivity=0.0)
    spec.ding = replace(spec.ding, gas_production_rate=0.0, temperature_enabled=False)
    spec.initial_velocity = (0.2, -0.1, 0.05)
    return spec


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_measure_closed_forms_on_an_insulated_translating_cube(precision):
    spec = _insulated_translating_cube()
    o = oracle_from_spec(spec, precision)
    V0 = np.asarray(spec.samples.reference_volumes, np.float64)
    t, c = spec.transport, spec.ding
    for _ in range(6):
        o.Step()
    m = o.Measure()
    np.testing.assert_allclose(m["particle_volume_m3"], V0.sum(), rtol=2e-6)
    np.testing.assert_allclose(m["temperature_mean_k"], t.initial_temperature_k, rtol=1e-12)
    np.testing.assert_allclose(m["water_mass_kg"], t.water_density * t.initial_saturation * V0.sum(), rtol=1e-12)
    moles = c.gas_initial_pressure_pa * c.porosity / (O.GAS_CONSTANT * t.initial_temperature_k)
    np.testing.assert_allclose(m["gas_total_mol"], moles * V0.sum(), rtol=1e-12)
    assert m["water_lost_kg"] == 0.0 and m["heat_exchange_j"] == 0.0
    cap = spec.material.density * t.specific_heat * V0
    np.testing.assert_allclose(m["thermal_energy_j"], t.initial_temperature_k * cap.sum(), rtol=1e-12)
    # a planted temperature field: the capacity-weighted mean (equal capacities here -> the plain mean of a
    # field whose two halves sit at T1 -/+ 10 K: T1 exactly)
    T = np.full(o.n, t.initial_temperature_k)
    T[: o.n // 2] -= 10.0
    T[o.n - o.n // 2:] += 10.0
    o.state.transport["T"][...] = T
    np.testing.assert_allclose(o.Measure()["temperature_mean_k"], t.initial_temperature_k, rtol=1e-12)
    o.state.transport["T"][...] = 0.0
    o.state.transport["T"][0] = 1.0
    np.testing.assert_allclose(o.Measure()["temperature_mean_k"], cap[0] / cap.sum(), rtol=1e-12)


def test_measure_volume_tracks_a_uniform_dilation():
    """sum V0 det F for F = (1 + e) I planted on every particle: (1 + e)^3 sum V0."""
    spec = _insulated_translating_cube()
    o = oracle_from_spec(spec, "f64")
    e = 0.01
    o.state.F[...] = (1.0 + e) * np.eye(3)
    V0 = np.asarray(spec.samples.reference_volumes, np.float64)
    np.testing.assert_allclose(o.Measure()["particle_volume_m3"], (1 + e) ** 3 * V0.sum(), rtol=1e-13)


def test_sphere_fixture_is_the_oracle_run():
    """The first 100 substeps of the committed 1000-substep fixture, re-derived by both oracles."""
    import sys
    sys.path.insert(0, str(GOLDEN))
    import make_sphere_small_1000 as gen
    data = np.load(GOLDEN / "sphere_small_1000.npz")
    assert tuple(data["keys"]) == gen.KEYS and int(data["steps"]) == 1000 and int(data["every"]) == gen.EVERY
    for precision in ("f32", "f64"):
        _, rows = gen.run(precision, steps=100)
        np.testing.assert_allclose(rows, data[f"aggregates_{precision}"][: len(rows)], rtol=1e-12, atol=0)
    # the recorded FP32 and FP64 runs are the same physics: aggregates within 1e-6 relative along the run
    a32, a64 = data["aggregates_f32"], data["aggregates_f64"]
    main = [gen.KEYS.index(k) for k in ("particle_volume_m3", "temperature_mean_k", "water_mass_kg", "gas_total_mol")]
    np.testing.assert_allclose(a32[:, main], a64[:, main], rtol=1e-6)


def test_gradient_fixture_loss_is_the_oracle_run():
    """The FP64 loss recorded in ``tests/golden/sphere_gradient_fd.npz`` re-derived (one 256-substep run), and the
    loss function itself: numpy and torch forms agree, and it vanishes at its targets."""
    import sys
    import torch
    from baking_amd import scenes
    sys.path.insert(0, str(GOLDEN))
    import make_sphere_gradient_fd as gen
    data = np.load(GOLDEN / "sphere_gradient_fd.npz")
    np.testing.assert_allclose(gen.loss_of(("base", -1, 0.0, "f64")), float(data["loss_f64"]), rtol=1e-12)
    spec = scenes.sphere_gradient()
    rng = np.random.default_rng(0)
    n = len(spec.samples.particle_positions)
    x = spec.samples.particle_positions + rng.normal(0, 1e-3, (n, 3))
    W = np.full(n, spec.transport.water_density * spec.transport.initial_saturation * spec.samples.reference_volumes[0])
    F = np.eye(3) + rng.normal(0, 1e-2, (n, 3, 3))
    a = scenes.gradient_loss(spec, x, W, F)
    b = scenes.gradient_loss(spec, torch.tensor(x), torch.tensor(W), torch.tensor(F), xp=torch)
    np.testing.assert_allclose(float(b), a, rtol=1e-12)
    # at the targets: a rigid shift of the centroid onto y* and W scaled onto s* make L = 0
    t = spec.transport
    V0 = np.asarray(spec.samples.reference_volumes)
    dry = (spec.material.density - t.water_density * t.initial_saturation) * V0
    W1 = t.water_density * spec.notes["saturation_target"] * V0
    m = dry + W1
    x1 = spec.samples.particle_positions.copy()
    x1[:, 1] += spec.notes["centroid_target"] - np.dot(m, x1[:, 1]) / m.sum()
    assert scenes.gradient_loss(spec, x1, W1, np.tile(np.eye(3), (n, 1, 1))) < 1e-28
