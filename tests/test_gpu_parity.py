"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Cell keys and the sort permutation are bit-exact against the oracle's keys
and numpy's stable argsort; per-substep particle and grid fields agree within
1e-5 relative (scale-relative max error, FP32); reference closed forms hold.
"""
import numpy as np
import pytest

from tests.harness import O, check_fields, gpu_fields, oracle_from_spec, three_way

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import _native, scenes  # noqa: E402

RTOL = 1.0e-5


def _sorted_on_gpu(keys_np, key_bits, n_cap=None):
    import ctypes
    lib = _native.load()
    n = len(keys_np)
    ws = lib.bk_workspace_create(max(n, 1), 8, 0, 0)
    assert ws
    try:
        kin = torch.as_tensor(keys_np.astype(np.int64).astype(np.int32), device="cuda")
        kout = torch.empty_like(kin)
        vout = torch.empty(n, dtype=torch.int32, device="cuda")
        _native.check(lib.bk_sort_pairs(ctypes.c_void_p(ws), _native.ptr(kin), None, _native.ptr(kout), _native.ptr(vout), n,
                                        key_bits, _native.stream_ptr()), "bk_sort_pairs")
        torch.cuda.synchronize()
        return kout.cpu().numpy().view(np.uint32), vout.cpu().numpy()
    finally:
        lib.bk_workspace_destroy(ctypes.c_void_p(ws))


@pytest.mark.parametrize("n,key_bits,seed", [(1, 8, 0), (2047, 8, 1), (2049, 12, 2), (100000, 24, 3), (300001, 18, 4)])
def test_radix_sort_bit_exact(n, key_bits, seed):
    rng = np.random.default_rng(seed)
    # many duplicates so stability matters
    keys = rng.integers(0, min(1 << key_bits, max(n // 7, 2)), size=n).astype(np.uint32)
    sk, perm = _sorted_on_gpu(keys, key_bits)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(perm, ref)
    np.testing.assert_array_equal(sk, keys[ref])


@pytest.mark.parametrize("tiled", [0, 1])
def test_cell_keys_bit_exact(tiled):
    spec = scenes.sphere_small()
    x = spec.samples.particle_positions.astype(np.float32)
    # move some particles onto exact cell boundaries and half-cells
    dx = np.float32(spec.simulation.grid_spacing)
    x[:64] = (np.floor(x[:64] / dx) * dx).astype(np.float32)
    x[64:128] = ((np.floor(x[64:128] / dx) + np.float32(0.5)) * dx).astype(np.float32)
    ref, ref_flags = O.cell_keys(x, spec.simulation.grid_origin, spec.simulation.grid_spacing,
                                 spec.simulation.grid_resolution, tiled=bool(tiled))
    lib = _native.load()
    xt = torch.as_tensor(x, device="cuda")
    keys = torch.empty(len(x), dtype=torch.int32, device="cuda")
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    o, keep = _native.host_f32(spec.simulation.grid_origin)
    _native.check(lib.bk_cell_keys(_native.ptr(xt), len(x), o, spec.simulation.grid_spacing, spec.simulation.grid_resolution,
                                   tiled, _native.ptr(keys), _native.ptr(flags), _native.stream_ptr()), "bk_cell_keys")
    got = keys.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, ref)
    assert flags.cpu().numpy()[1] == ref_flags[1] == 0
    sk, perm = _sorted_on_gpu(got, 24)
    np.testing.assert_array_equal(perm, O.sort_permutation(ref))


def test_sphere_substep_parity():
    """The small plain scene (Ding + gas + heat/water transport): every field, 3 then 20 substeps."""
    three_way(scenes.sphere_small(), 20, rtol=RTOL, check_every=1)


@pytest.mark.parametrize("solver", ["bracketed", "bisection"])
@pytest.mark.parametrize("builder,steps", [("demo_stretch", 200), ("demo_gas", 40), ("demo_heated", 40)])
def test_loaded_dough_parity(builder, steps, solver):
    """Plastic flow, gas-driven porous expansion and heated porous flow: the return map under load,
    with both coupled-porosity solvers against the oracle's literal bisection."""
    spec = getattr(scenes, builder)()
    three_way(spec, steps, rtol=RTOL, check_every=10, solver=solver)


@pytest.mark.parametrize("solver", ["bisection", "bracketed"])
def test_solve_porosity_matches_oracle(solver):
    """Coupled porosity on random and realistic inputs, plus the near-closure known answer
    (reference tests/test_gas.py:208-222): float32 results equal the oracle's bisection."""
    from baking_amd.physics.materials.porous_return import solve_porosity
    rng = np.random.default_rng(11)
    m = 4000
    por = rng.uniform(0.02, 0.08, m)
    cols = [rng.uniform(0.5, 20, m), rng.uniform(0, 30, m), 10 ** rng.uniform(-4, 1, m), 10 ** rng.uniform(-4, 1, m),
            rng.uniform(-0.1, 0.05, m)]
    real = [rng.uniform(1.0, 1.5, m), rng.uniform(0, 0.05, m), 10 ** rng.uniform(-1, 1, m), 10 ** rng.uniform(-3, 0, m),
            rng.uniform(-0.01, 0.003, m)]
    q, p, d, h, ex = [np.concatenate([a, b]).astype(np.float32) for a, b in zip(cols, real)]
    n0 = np.float32(0.05)
    pv = 1.0 - float(n0) + np.concatenate([por, por]).astype(np.float32).astype(np.float64)
    q = np.append(q, np.float32(4.115)); p = np.append(p, np.float32(29.464)); d = np.append(d, np.float32(0.001052))
    h = np.append(h, np.float32(0.003390)); ex = np.append(ex, np.float32(-0.0230067))
    pv = np.append(pv, 1.0 - float(np.float32(0.2)) + float(np.float32(3.725290298461914e-7)))
    n0s = np.full(len(q), n0, np.float32)
    n0s[-1] = np.float32(0.2)
    got, pred = solve_porosity(torch.tensor(q), torch.tensor(p), torch.tensor(d), torch.tensor(h), torch.tensor(ex),
                               torch.tensor(pv), torch.tensor(n0s), solver=solver)
    got = got.cpu().numpy()
    ref = np.array([O.solve_porosity(float(a), float(b), float(c), float(e), float(f), float(g), float(k))
                    for a, b, c, e, f, g, k in zip(q, p, d, h, ex, pv, n0s)], dtype=np.float32)
    assert 0 < got[-1] < 3.725290298461914e-7
    np.testing.assert_allclose(got[-1], pred.cpu().numpy()[-1], atol=1e-12)
    body = slice(0, len(q) - 1)
    # fp64 log/exp of the device and of the host libm may differ in the last bit, which can flip a
    # bisection predicate evaluated within ~1e-16 of the root: allow a couple of 1-ulp float differences.
    mismatch = np.flatnonzero(got[body] != ref[body])
    print(solver, "float mismatches vs the oracle bisection:", len(mismatch), "of", len(ref) - 1)
    assert len(mismatch) <= 2, f"{len(mismatch)} float mismatches vs the reference bisection"
    np.testing.assert_allclose(got[body], ref[body], rtol=1e-6)


@pytest.mark.parametrize("resolution,samples_per_tet", [(2, 150), (3, 40), (2, 1)])
def test_mesh_reconstruction_matches_oracle(resolution, samples_per_tet):
    """geometry/mesh_tracking.py:13-38 on seeded perturbed particles: 300-3600 bindings per vertex (several strides of
    a 256-thread workgroup, a partial last stride) and 1-24.  The GPU sums each vertex's bindings in another order than
    the oracle's loop, so it is held to the FP64 oracle at a few FP32 roundings."""
    from baking_amd.geometry.mesh_loader import CreateDemoMesh
    from baking_amd.geometry.mesh_tracking import BuildBinding, ReconstructMeshVertices
    from baking_amd.geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(resolution)
    samples = SampleParticles(mesh, 1000.0, samples_per_tet)
    b = BuildBinding(mesh, samples)
    rng = np.random.default_rng(resolution * 1000 + samples_per_tet)
    n, nv = len(samples.particle_positions), len(mesh.vertex_positions)
    x = (samples.particle_positions + rng.normal(0.0, 1e-3, (n, 3))).astype(np.float32)
    F = (np.eye(3)[None] + rng.normal(0.0, 0.05, (n, 3, 3))).astype(np.float32)
    rv = mesh.vertex_positions.astype(np.float32)
    host = (x, F, rv, b["offsets"], b["particles"], b["weights"], b["reference_offsets"])
    out = torch.empty(nv, 3, device="cuda")
    ReconstructMeshVertices(nv, *[torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in host], out)
    got = out.cpu().numpy()
    want = {}
    for prec in ("f32", "f64"):
        R = np.float32 if prec == "f32" else np.float64
        o = np.zeros((nv, 3), R)
        O.Kernels(prec).ReconstructMeshVertices(nv, *[np.ascontiguousarray(a, R) if a.dtype.kind == "f" else a for a in host], o)
        want[prec] = o
    err, oerr = np.abs(got - want["f64"]).max(), np.abs(want["f32"] - want["f64"]).max()
    print(f"res {resolution} x{samples_per_tet}: |gpu - f64| {err:.2e}, |f32 oracle - f64| {oerr:.2e}", flush=True)
    # the FP32 oracle is within ~4e-8 of FP64; one dropped binding would move a vertex by ~1e-6
    np.testing.assert_allclose(got, want["f64"], atol=2e-7)


def test_free_fall_symplectic_euler():
    """tests/test_mpm.py:40-53 of the reference on the GPU path."""
    from dataclasses import replace
    from baking_amd.config import MaterialConfig, SimulationConfig
    from baking_amd.geometry.mesh_loader import CreateDemoMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    from baking_amd.physics.mpm import MpmSimulator
    mesh = CreateDemoMesh(3)
    cfg = replace(SimulationConfig(), device_name="auto", gravity=(0.0, -9.81, 0.0))
    sim = MpmSimulator(mesh, SampleParticles(mesh, 1000.0, 4), cfg, MaterialConfig(), (0.0, 0.0, 0.0))
    x0 = sim.current_state.particle_positions.cpu().numpy().copy()
    n, dt = 40, cfg.time_step
    sim.Advance(n)
    exp = x0.copy()
    exp[:, 1] -= 9.81 * dt ** 2 * n * (n + 1) / 2
    np.testing.assert_allclose(sim.current_state.particle_positions.cpu().numpy(), exp, atol=1.5e-6)
    v = np.zeros_like(x0)
    v[:, 1] = -9.81 * dt * n
    np.testing.assert_allclose(sim.current_state.particle_velocities.cpu().numpy(), v, atol=1.0e-5)


def test_free_translation_conserves_mass_and_shape():
    """tests/test_mpm.py:18-37 of the reference on the GPU path."""
    from dataclasses import replace
    from baking_amd.config import MaterialConfig, SimulationConfig
    from baking_amd.geometry.mesh_loader import CreateDemoMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    from baking_amd.physics.mpm import MpmSimulator
    mesh = CreateDemoMesh(3)
    cfg = replace(SimulationConfig(), device_name="auto", gravity=(0.0, 0.0, 0.0))
    v0 = (0.2, -0.1, 0.05)
    sim = MpmSimulator(mesh, SampleParticles(mesh, 1000.0, 4), cfg, MaterialConfig(), v0)
    x0 = sim.current_state.particle_positions.cpu().numpy().copy()
    verts0 = sim.vertex_positions.cpu().numpy().copy()
    m0 = sim.particle_data.particle_masses.cpu().numpy().sum(dtype=np.float64)
    stats = sim.Advance(30)
    t = 30 * cfg.time_step
    np.testing.assert_allclose(sim.current_state.particle_positions.cpu().numpy(), x0 + t * np.array(v0), atol=2.0e-6)
    np.testing.assert_allclose(sim.current_state.particle_velocities.cpu().numpy(), np.tile(v0, (len(x0), 1)), atol=2.0e-5)
    np.testing.assert_allclose(sim.current_state.deformation_gradients.cpu().numpy(), np.tile(np.eye(3), (len(x0), 1, 1)),
                               atol=1.0e-5)
    np.testing.assert_allclose(sim.vertex_positions.cpu().numpy(), verts0 + t * np.array(v0), atol=2.0e-6)
    np.testing.assert_allclose(stats["grid_mass"], m0, rtol=1.0e-6)
    assert stats["unmapped_mesh_vertices"] == 0


def test_far_off_and_nonfinite_particles_are_flagged_not_indexed():
    """Positions whose cell base overflows an int, or that are not finite, fail the stencil test in every
    kernel (flag 1), so no node index is formed from them: the substep completes and the flag is raised."""
    import ctypes
    spec = scenes.sphere_small()
    sim = scenes.build_simulator(spec)
    x = sim.current_state.particle_positions
    x[0, 0] = 1.0e30
    x[1, 1] = float("inf")
    x[2, 2] = float("nan")
    x[3, 0] = -1.0e30
    sim._launch_substeps(2)  # raw substeps: Advance would raise the reference's errors for the flags
    torch.cuda.synchronize()
    assert sim._read_flags()[1] == 1
    lib = _native.load()
    keys = torch.empty(4, dtype=torch.int32, device="cuda")
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    o, keep = _native.host_f32(spec.simulation.grid_origin)
    bad = torch.tensor([[1e30, 0.5, 0.5], [0.5, float("inf"), 0.5], [0.5, 0.5, float("nan")], [0.5, 0.5, 0.5]],
                       dtype=torch.float32, device="cuda")
    _native.check(lib.bk_cell_keys(_native.ptr(bad), 4, o, spec.simulation.grid_spacing, spec.simulation.grid_resolution, 1,
                                   _native.ptr(keys), _native.ptr(flags), _native.stream_ptr()), "bk_cell_keys")
    assert flags.cpu().numpy()[1] == 1
    assert keys.cpu().numpy()[:3].tolist() == [0, 0, 0]
    del ctypes


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [-1, 1])
def test_cube_stops_at_each_wall(axis, side):
    """The fused substep's wall projection: an elastic cube fired at 1 m/s into each of the six walls (reference
    physics/grid.py:32-37, tests/test_sphere.py:44-47).  Every occupied node of the wall's layers moves inward or
    tangentially, the cube rebounds without leaving the valid stencil region, and every field agrees with the
    FP32/FP64 oracles after each 300 substeps."""
    from tests.test_oracle_walls import band_violation, wall_scene
    spec = wall_scene(axis, side)
    sim_cfg = spec.simulation
    res, bc = sim_cfg.grid_resolution, sim_cfg.boundary_cells
    gpu = scenes.build_simulator(spec)
    o32, o64 = oracle_from_spec(spec, "f32"), oracle_from_spec(spec, "f64")
    mass = o32.m0.astype(np.float64)
    p0 = float(np.dot(mass, o32.state.v[:, axis])) * side
    for step in range(1, spec.steps + 1):
        gpu.Step()
        o32.Step()
        o64.Step()
        if step % 300 == 0:
            check_fields(f"{spec.name} substep {step}", gpu_fields(gpu), o32, o64)
            nm = gpu.grid_state.node_masses.cpu().numpy()
            nv = gpu.grid_state.node_velocities.cpu().numpy()
            assert band_violation(nm, nv, res, bc, axis, side) == 0.0
    v = gpu.current_state.particle_velocities.cpu().numpy().astype(np.float64)
    assert float(np.dot(mass, v[:, axis])) * side < -0.1 * p0
    assert gpu.CheckState()["minimum_j"] > 0


def _substep_order(sim, n):
    """The last substep's keys (particle order), sorted keys and permutation, read from its workspace."""
    lib = _native.load()
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    sorted_keys = torch.empty_like(keys)
    order = torch.empty_like(keys)
    _native.check(lib.bk_workspace_sort_order(sim._workspace, _native.ptr(keys), _native.ptr(sorted_keys), _native.ptr(order), n,
                                              _native.stream_ptr()), "bk_workspace_sort_order")
    torch.cuda.synchronize()
    return (keys.cpu().numpy().view(np.uint32), sorted_keys.cpu().numpy().view(np.uint32), order.cpu().numpy())


def _expected_keys(spec, x, block_offset=0):
    sim = spec.simulation
    keys, flags = O.cell_keys(x, sim.grid_origin, sim.grid_spacing, sim.grid_resolution, tiled=True)
    assert flags[1] == 0
    return keys + np.uint32(block_offset << 6)


def _check_order(keys, sorted_keys, order, want_keys):
    np.testing.assert_array_equal(keys, want_keys)
    np.testing.assert_array_equal(order, np.argsort(want_keys, kind="stable"))
    np.testing.assert_array_equal(sorted_keys, want_keys[order])


def test_substep_sort_is_the_stable_sort_of_the_keys():
    """The fused substep's particle order (reference key physics/transfer.py:16-33 on the tiled grid) equals
    numpy's stable argsort of the oracle's keys, bit for bit, on the small sphere after 1 and 5 substeps."""
    spec = scenes.sphere_small()
    sim = scenes.build_simulator(spec)
    n = sim.particle_data.particle_count
    for steps in (1, 4):
        x = sim.current_state.particle_positions.cpu().numpy().copy()
        sim._launch_substeps(1)
        _check_order(*_substep_order(sim, n), _expected_keys(spec, x))
        if steps > 1:
            sim._launch_substeps(steps - 1)


def test_batched_substep_sort_is_the_stable_sort_of_the_keys():
    from baking_amd.physics.batch import BatchSimulator
    specs = scenes.batched_doughs(count=3, seed=8, half_extent=(0.02, 0.035), max_particles=6000)
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs])
    batch.Advance(2)
    xs = [d.current_state.particle_positions.cpu().numpy().copy() for d in batch.doughs]
    batch.Advance(1)
    want = np.concatenate([_expected_keys(s, x, int(batch.block_offsets[d])) for d, (s, x) in enumerate(zip(specs, xs))])
    _check_order(*_substep_order(batch, batch.particle_count), want)


def _crowded_spec():
    """A 0.06 m cube sampled 400 particles per tet: more than 1024 particles in one 4x4x4 block."""
    from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
    from baking_amd.geometry.mesh_loader import VoxelTetMesh
    from baking_amd.geometry.mesh_validation import ValidateTetMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    verts, tets = VoxelTetMesh((1, 1, 1), 0.06, (0.52, 0.52, 0.52))
    mesh = ValidateTetMesh(verts, tets)
    return scenes.SceneSpec("crowded", mesh, SampleParticles(mesh, 1000.0, 400), SimulationConfig(gravity=(0.0, -9.81, 0.0)),
                            MaterialConfig(), DingConfig(), TransportConfig(), steps=2)


def test_crowded_block_sort_is_the_stable_sort_of_the_keys():
    """More than 1024 particles in one 4x4x4 block (the sort's LDS capacity): the global-memory ranking path."""
    spec = _crowded_spec()
    sim = scenes.build_simulator(spec)
    n = sim.particle_data.particle_count
    x = sim.current_state.particle_positions.cpu().numpy().copy()
    sim._launch_substeps(1)
    want = _expected_keys(spec, x)
    assert np.bincount(want >> 6).max() > 1024
    _check_order(*_substep_order(sim, n), want)


def test_dense_tile_substep_parity():
    """A 0.06 m cube of 6^3 voxels, 2 particles per tet, of the transport demo's dough (Ding, pore gas, heat and
    water transport): all 2,592 particles and the 419 owners of its boundary faces fall in one 4x4x4 block.  So one
    P2G tile holds more than 1024 particles (the mechanical P2G fills and walks its (cell, rank) work list in several
    windows) and more face owners than the thermal P2G lists (it sweeps the tile for them).  Three substeps against
    the FP32/FP64 oracle."""
    from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
    from baking_amd.geometry.mesh_loader import VoxelTetMesh
    from baking_amd.geometry.mesh_validation import ValidateTetMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    from baking_amd.physics.thermal_water import BoundaryFaces
    verts, tets = VoxelTetMesh((6, 6, 6), 0.01, (0.52, 0.52, 0.52))
    mesh = ValidateTetMesh(verts, tets)
    samples = SampleParticles(mesh, 1000.0, 2)
    sim = SimulationConfig(time_step=1e-4, gravity=(0.0, -9.81, 0.0), grid_spacing=0.025, grid_resolution=40)
    ding = DingConfig(gas_enabled=True, porosity=0.2, gas_production_rate=0.001, temperature_enabled=True)
    spec = scenes.SceneSpec("dense_tile", mesh, samples, sim, MaterialConfig(model="ding2019", young_modulus=10000.0), ding,
                            TransportConfig(enabled=True, heat_transfer=20.0), steps=3)
    keys = _expected_keys(spec, samples.particle_positions.astype(np.float32))
    owners = np.unique(BoundaryFaces(mesh, samples)[1])
    assert np.bincount(keys >> 6).max() > 1024
    assert np.bincount(keys[owners] >> 6).max() > 128
    three_way(spec, 3, rtol=RTOL, check_every=1)


@pytest.mark.parametrize("sticky", [True, False])
def test_sticky_floor_on_the_gpu(sticky):
    """The fused substep's floor option (MpmSimulator.SetStickyFloor; no reference counterpart) against the oracle
    with the same option: a cube sliding at 0.5 m/s with its bottom just inside the simulator's initial domain
    (2.04 dx up, so its lowest particles reach the floor layers).  Sticky, every occupied floor-layer node has
    velocity 0 and the bottom particles are braked (below 0.2 m/s within 50 substeps); off, the reference's
    frictionless floor lets them slide at 0.5 m/s."""
    from tests.test_oracle_walls import sliding_cube
    spec = sliding_cube(sticky, bottom_cells=2.04)
    gpu = scenes.build_simulator(spec)
    gpu.SetStickyFloor(sticky)
    o32, o64 = oracle_from_spec(spec, "f32", sticky_floor=sticky), oracle_from_spec(spec, "f64", sticky_floor=sticky)
    sim = spec.simulation
    res, bc, dx = sim.grid_resolution, sim.boundary_cells, sim.grid_spacing
    low = (gpu.current_state.particle_positions.cpu().numpy()[:, 1] - sim.grid_origin[1]) / dx < 2.5
    assert low.sum() > 0
    for step in range(1, spec.steps + 1):
        gpu.Step()
        o32.Step()
        o64.Step()
        if step % 10 == 0:
            check_fields(f"{spec.name} substep {step}", gpu_fields(gpu), o32, o64)
            m = gpu.grid_state.node_masses.cpu().numpy().reshape((res,) * 3)
            v = gpu.grid_state.node_velocities.cpu().numpy().reshape((res,) * 3 + (3,))
            occupied = m[:, :bc] > 0
            assert occupied.any()
            if sticky:
                assert np.all(v[:, :bc][occupied] == 0.0)
    vx = gpu.current_state.particle_velocities.cpu().numpy()[low, 0]
    if sticky:
        assert vx.max() < 0.2
    else:
        assert np.all(np.abs(vx - 0.5) < 0.05)
