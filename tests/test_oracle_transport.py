"""Oracle pins for heat and water transport and mesh tracking.

Restates reference tests/test_transport.py and tests/test_mesh_reconstruction.py
against the CPU oracle: insulated invariance, water and energy balances,
boundary-face areas, stability bound, the cosine heat mode's convergence
orders and discrete eigenmode, and affine mesh reconstruction.
"""
from dataclasses import replace

import numpy as np
import pytest

from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
from baking_amd.geometry.mesh_loader import CreateDemoMesh
from baking_amd.geometry.mesh_validation import ValidateTetMesh
from baking_amd.geometry.particle_sampling import SampleParticles
from tests.harness import O


def build_transport(config=None, spacing=0.025, dt=1e-4, subdivisions=2, gas=False, precision="f32"):
    mesh = CreateDemoMesh(subdivisions)
    sim = replace(SimulationConfig(), device_name="cpu", time_step=dt, gravity=(0., 0., 0.), grid_spacing=spacing,
                  grid_resolution=int(round(1 / spacing)))
    ding = DingConfig(gas_enabled=gas, porosity=0.2 if gas else 0., gas_production_rate=0.001, temperature_enabled=True)
    return O.OracleSimulator(mesh, SampleParticles(mesh, 1000., 4), sim, MaterialConfig(model="ding2019", young_modulus=10000.),
                             (0.0, 0.0, 0.0), ding, config or TransportConfig(enabled=True), precision=precision)


def thermal_steps(sim, count, dt):
    for _ in range(count):
        sim.state = sim.transport_advance(sim.state, dt=dt)[0]
    return sim.state.transport


def test_insulated_uniform_temperature_and_water_conservation():
    """Reference tests/test_transport.py:30-47."""
    cfg = TransportConfig(enabled=True, heat_transfer=0., tray_heat_transfer=0., surface_water_loss=0.)
    sim = build_transport(cfg)
    w0 = sim.state.transport["W"].sum()
    t = thermal_steps(sim, 8, 0.001)
    np.testing.assert_allclose(t["T"], cfg.initial_temperature_k, atol=1e-12)
    np.testing.assert_allclose(t["W"].sum(), w0, rtol=2e-7)
    masses = t["W"] * np.where(sim.state.x[:, 0] < 0.5, 1.5, 0.5)
    sim.state.transport["W"] = masses.copy()
    after = thermal_steps(sim, 8, 0.01)["W"]
    np.testing.assert_allclose(after.sum(), masses.sum(), rtol=2e-7)
    high = masses > masses.mean()
    assert (after[high] - masses[high]).sum() < 0


def test_boundary_area_tray_and_water_stress():
    """Reference tests/test_transport.py:50-65: 48 faces, exposed 0.45, contact 0.09, water stress 1000 I."""
    sim = build_transport()
    assert len(sim.faces) == 48
    k = sim.k
    ex, co = np.zeros(sim.n), np.zeros(sim.n)
    k.SurfaceAreas(len(sim.faces), sim.reference_vertices, sim.faces, sim.owners, ex, co, .45, .001)
    np.testing.assert_allclose(ex.sum(), .45, rtol=2e-6)
    np.testing.assert_allclose(co.sum(), .09, rtol=2e-6)
    before = np.zeros((sim.n, 3, 3), np.float32)
    after = np.zeros_like(before)
    k.WaterStress(sim.n, sim.state.F, np.full(sim.n, .3), before, after, .4, 10000.)
    np.testing.assert_allclose(after, np.tile(np.eye(3) * 1000, (sim.n, 1, 1)), rtol=1e-6)


def test_outside_in_water_loss_heat_balance_and_latent_cooling():
    """Reference tests/test_transport.py:68-87."""
    common = TransportConfig(enabled=True, conductivity=20., specific_heat=100., heat_transfer=2000., surface_water_loss=1e-5,
                             water_diffusivity=1e-8, water_stress=0.)
    means = []
    for latent in (0., 2.26e6):
        sim = build_transport(replace(common, latent_heat=latent), subdivisions=3)
        t = thermal_steps(sim, 40, .001)
        T = t["T"]
        p = sim.samples.particle_positions
        surface = np.zeros(len(T), bool)
        surface[sim.owners] = True
        core = np.linalg.norm(p - np.array([.5, .6, .5]), axis=1) < .07
        assert T[surface].mean() > T[core].mean()
        w_init = common.water_density * common.initial_saturation * .027
        assert t["W"].sum() < w_init
        np.testing.assert_allclose(t["W"].sum() + t["losses"].sum(), w_init, rtol=2e-7)
        np.testing.assert_allclose(np.dot(sim.capacities, T - common.initial_temperature_k), t["exchanges"].sum(), rtol=2e-6)
        means.append(T.mean())
    assert means[1] < means[0]


def test_explicit_transport_bound():
    """Reference tests/test_transport.py:174-178."""
    sim = build_transport(TransportConfig(enabled=True, conductivity=1e8))
    with pytest.raises(RuntimeError, match="stability bound"):
        sim.transport_advance(sim.state, dt=.1)


def grid_decay(resolution, dt):
    """Insulated cosine heat mode T = 300 + 10 cos(pi x) exp(-D pi^2 t) through DiffuseGrid."""
    k = O.Kernels("f32")
    count = resolution ** 3
    spacing = 1. / resolution
    coords = np.indices((resolution,) * 3)[0].reshape(-1)
    initial = 300 + 10 * np.cos(np.pi * (coords + .5) * spacing)
    T = initial.copy()
    vol = np.full(count, spacing ** 3)
    cap = np.full(count, 1000 * spacing ** 3)
    zero = np.zeros(count)
    flags = np.zeros(2, np.int32)
    for _ in range(round(.2 / dt)):
        dT, dW = np.zeros(count), np.zeros(count)
        k.DiffuseGrid(O.i64(count), vol, cap, T, zero, zero, zero, zero, zero, dT, dW, flags, 10., 0., 1000., 0, spacing,
                      resolution, dt)
        T = T + dT
    exact = 300 + (initial - 300) * np.exp(-.01 * np.pi ** 2 * .2)
    assert flags.max() == 0
    amplitude = np.dot(T - 300, initial - 300) / np.dot(initial - 300, initial - 300)
    return float(np.sqrt(np.mean((T - exact) ** 2))), float(amplitude)


def test_grid_and_time_convergence():
    """Reference tests/test_transport.py:181-213: second order in space, first order in time vs the discrete eigenmode."""
    errors = [grid_decay(n, .0005)[0] for n in (8, 16, 32)]
    assert errors[1] < .35 * errors[0] and errors[2] < .35 * errors[1]
    n = 16
    rate = 4 * .01 * n * n * np.sin(np.pi / (2 * n)) ** 2
    exact = np.exp(-rate * .2)
    time_errors = [abs(grid_decay(n, dt)[1] - exact) for dt in (.02, .01, .005)]
    assert time_errors[1] < .51 * time_errors[0] and time_errors[2] < .51 * time_errors[1]


def test_local_temperature_changes_gas_and_plastic_parameters():
    """Reference tests/test_transport.py:124-147 (one full substep with per-particle temperatures)."""
    cfg = TransportConfig(enabled=True, conductivity=0., heat_transfer=0., tray_heat_transfer=0., water_diffusivity=0.,
                          surface_water_loss=0., water_stress=0.)
    sim = build_transport(cfg, gas=True)
    values = np.linspace(300., 360., sim.n)
    sim.state.transport["T"][...] = values
    before = sim.state.ding["moles"].copy()
    sim.Step()
    d = sim.state.ding
    ding = sim.ding_cfg
    expected = before + sim.dt * 1000 * ding.gas_production_rate * np.exp(
        -((values - ding.gas_reaction_temperature_k) / ding.gas_reaction_width_k) ** 2)
    np.testing.assert_allclose(d["moles"] - before, expected - before, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(d["T"], values, atol=2e-5)
    yields = [ding.yield_stress * np.interp(t, ding.temperature_knots_k, ding.yield_factors) for t in values]
    np.testing.assert_allclose(d["yield_stress"], yields, rtol=1e-5)
    np.testing.assert_array_equal(d["gelled"], values >= ding.temperature_knots_k[-1])
    J = np.linalg.det(sim.state.F)
    np.testing.assert_allclose(d["pressure"], expected * O.GAS_CONSTANT * d["T"] / (J * d["porosity"]), rtol=5e-7)


@pytest.mark.parametrize("sample_count", [1, 4, 7])
def test_reconstruction_reproduces_affine_deformation(sample_count):
    """Reference tests/test_mesh_reconstruction.py:15-41."""
    mesh = CreateDemoMesh(2)
    samples = SampleParticles(mesh, 1000.0, sample_count)
    b = O.mesh_binding(mesh.vertex_positions, mesh.tet_indices, samples.source_tets, samples.particle_positions,
                       samples.reference_volumes)
    k = O.Kernels("f32")
    angle = 0.7
    R = np.array([[np.cos(angle), -np.sin(angle), 0.], [np.sin(angle), np.cos(angle), 0.], [0., 0., 1.]])
    t = np.array([0.15, -0.1, 0.05])
    for A in (np.eye(3), R, R @ np.array([[1.2, 0.2, 0.0], [0.0, 0.8, 0.1], [0.0, 0.0, 1.1]])):
        x = (samples.particle_positions @ A.T + t).astype(np.float32)
        F = np.tile(A.astype(np.float32), (len(x), 1, 1))
        out = np.zeros((len(mesh.vertex_positions), 3), np.float32)
        k.ReconstructMeshVertices(len(out), x, F, mesh.vertex_positions.astype(np.float32), b["offsets"], b["particles"],
                                  b["weights"].astype(np.float32), b["reference_offsets"].astype(np.float32), out)
        np.testing.assert_allclose(out, mesh.vertex_positions @ A.T + t, rtol=1e-6, atol=4e-7)


def test_shared_vertices_blend_incident_particle_predictions():
    """Reference tests/test_mesh_reconstruction.py:44-74."""
    mesh = ValidateTetMesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -2], [5, 5, 5]], dtype=float),
                           np.array([[0, 1, 2, 3], [0, 2, 1, 4]]))
    samples = SampleParticles(mesh, 1000.0, 1)
    b = O.mesh_binding(mesh.vertex_positions, mesh.tet_indices, samples.source_tets, samples.particle_positions,
                       samples.reference_volumes)
    pos = samples.particle_positions.copy()
    pos[0] += [0.1, 0.0, 0.0]
    pos[1] += [0.0, 0.2, 0.0]
    F = np.array([np.diag([1.1, 0.9, 1.0]), np.diag([0.8, 1.2, 1.0])])
    out = np.zeros((6, 3), np.float32)
    O.Kernels("f32").ReconstructMeshVertices(6, pos.astype(np.float32), F.astype(np.float32),
                                             mesh.vertex_positions.astype(np.float32), b["offsets"], b["particles"],
                                             b["weights"].astype(np.float32), b["reference_offsets"].astype(np.float32), out)
    expected = mesh.vertex_positions.copy()
    for v in range(6):
        inc = [p for p, t in enumerate(samples.source_tets) if v in mesh.tet_indices[t]]
        if inc:
            preds = [pos[p] + F[p] @ (mesh.vertex_positions[v] - samples.particle_positions[p]) for p in inc]
            expected[v] = np.average(preds, axis=0, weights=samples.reference_volumes[inc])
    np.testing.assert_allclose(out, expected, atol=3e-7)
    assert b["unmapped"] == 1
    np.testing.assert_array_equal(out[5], mesh.vertex_positions[5].astype(np.float32))
