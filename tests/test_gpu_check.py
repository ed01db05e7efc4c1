"""GPU CheckState (reference physics/mpm.py:243-308): one fused reduction instead of host copies of the state.

The statistics must equal the oracle's numpy CheckState (the reference's host
computation) evaluated on the GPU simulator's own state: sums to FP64
rounding, float extrema (det F in float, by cofactors here and by LU in
numpy) to FP32 rounding.  Each error is planted in the state (a NaN, an
inverted F, a particle outside the stencil, a CFL violation, an unconverged or
nonfinite Ding state) and must raise the reference's text, computed by the
oracle on the same state, without updating the mesh.  The translating cube also
checks the reference's own assertion grid_mass = initial mass
(tests/test_mpm.py:36) and the closed forms J = 1, KE = M |v|^2 / 2.
"""
import numpy as np
import pytest

from tests.harness import oracle_on_gpu_state as _oracle_on

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import scenes  # noqa: E402
from baking_amd.physics.batch import BatchSimulator  # noqa: E402

_EXACT = ("step_count", "particle_count", "unmapped_mesh_vertices", "inverted_mesh_tets")
_SUMS = ("particle_volume_m3", "total_mass", "grid_mass", "kinetic_energy", "simulation_time")
_FLOATS = ("minimum_j", "maximum_j", "maximum_speed", "estimated_cfl")


def _compare(tag, got, want):
    assert set(got) == set(want), tag
    for k in _EXACT:
        assert got[k] == want[k], (tag, k, got[k], want[k])
    for k in _SUMS:  # particle volume: V0 in float here, in double in the reference's samples
        np.testing.assert_allclose(got[k], want[k], rtol=2e-7, err_msg=f"{tag}: {k}")
    for k in _FLOATS:
        np.testing.assert_allclose(got[k], want[k], rtol=1e-6, err_msg=f"{tag}: {k}")
    print(f"{tag}: " + ", ".join(f"{k} {got[k]:.9g}" for k in _SUMS + _FLOATS), flush=True)


@pytest.mark.parametrize("name,steps", [("demo_translation", 30), ("sphere_small", 12), ("demo_transport", 8), ("demo_resting", 6)])
def test_check_state_matches_host_numpy(name, steps):
    spec = getattr(scenes, name)()
    sim = scenes.build_simulator(spec)
    sim.Advance(steps, check=False)
    got = sim.CheckState()
    _compare(name, got, _oracle_on(spec, sim).check_state())
    assert sim.mesh_step_count == sim.step_count


def test_free_translation_closed_forms():
    """tests/test_mpm.py:18-37: grid_mass equals the initial mass; a rigid translation keeps J = 1 and KE = M|v|^2/2."""
    spec = scenes.demo_translation()
    sim = scenes.build_simulator(spec)
    mass = float(sim.particle_data.particle_masses.double().sum())
    stats = sim.Advance(30)
    v0 = np.array(spec.initial_velocity)
    np.testing.assert_allclose(stats["grid_mass"], mass, rtol=1e-6)
    np.testing.assert_allclose([stats["minimum_j"], stats["maximum_j"]], 1.0, atol=1e-5)
    np.testing.assert_allclose(stats["kinetic_energy"], 0.5 * mass * v0 @ v0, rtol=1e-4)
    np.testing.assert_allclose(stats["particle_volume_m3"], np.sum(spec.samples.reference_volumes), rtol=1e-5)
    assert stats["unmapped_mesh_vertices"] == 0 and stats["inverted_mesh_tets"] == 0
    v = sim.current_state.particle_velocities
    np.testing.assert_allclose(stats["estimated_cfl"], (float(v.norm(dim=1).max()) + spec.material.wave_speed)
                               * spec.simulation.time_step / spec.simulation.grid_spacing, rtol=1e-6)


def _plant(sim, what):
    s = sim.current_state
    p = 5
    if what == "nan_position":
        s.particle_positions[p, 0] = float("nan")
    elif what == "inverted":
        s.deformation_gradients[p] = torch.diag(torch.tensor([-1.0, 1.0, 1.0], device="cuda"))
    elif what == "outside_stencil":
        s.particle_positions[p, 1] = sim.simulation_config.grid_origin[1] + 0.2 * sim.simulation_config.grid_spacing
    elif what == "cfl":
        s.particle_velocities[p] = torch.tensor([400.0, 0.0, 0.0], device="cuda")
    elif what == "residual":
        s.material_state.return_residuals[p] = 1e-3
    elif what == "nonfinite_ding":
        s.material_state.elastic_gradients[p, 0, 0] = float("inf")


@pytest.mark.parametrize("what", ["nan_position", "inverted", "outside_stencil", "cfl", "residual", "nonfinite_ding"])
def test_check_state_errors(what):
    spec = scenes.sphere_small()
    sim = scenes.build_simulator(spec)
    sim.Advance(3, check=False)
    sim.UpdateMesh()
    verts = sim.vertex_positions.clone()
    _plant(sim, what)
    with pytest.raises(RuntimeError) as want:
        _oracle_on(spec, sim).check_state()
    with pytest.raises(RuntimeError) as got:
        sim.CheckState()
    print(what, "->", got.value, flush=True)
    assert str(got.value) == str(want.value)
    assert torch.equal(sim.vertex_positions, verts)  # the mesh is reconstructed only after every check


def test_check_state_of_a_batched_dough():
    specs = [scenes.demo_transport(), scenes.demo_translation()]
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs])
    batch.Advance(6)
    for d, spec in enumerate(specs):
        _compare(f"batched {spec.name}", batch.CheckState(d), _oracle_on(spec, batch.doughs[d]).check_state())
