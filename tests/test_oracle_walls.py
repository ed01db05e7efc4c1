"""The grid-wall projection of ``UpdateGrid`` (reference physics/grid.py:20-38) pinned without restating it.

* Projection: on random node masses and momenta (some empty nodes), the oracle's node velocity v at every
  occupied node is the nearest point to the free velocity u = p/m + dt g that satisfies the wall
  constraints, found here as the minimiser of |v - u| over a brute-force set of candidates per axis: on a
  node within ``boundary_cells`` of a low wall v_a >= 0, of a high wall v_a <= 0 (frictionless, unilateral:
  inward motion and tangential motion are left alone); empty nodes stay 0.
* Impact on each of the six walls: an elastic cube fired at 1 m/s into one wall stops and rebounds (its
  momentum into the wall changes sign within 900 substeps) and stays inside the grid's valid stencil region;
  every occupied node of that wall's layers moves inward or tangentially at every checked substep (the
  reference's floor check tests/test_sphere.py:44-47, v_y >= 0 in the floor layers, applied to every wall).
"""
import numpy as np
import pytest

from tests.harness import O, oracle_from_spec


def _nearest_admissible(u, low, high):
    """argmin over v of |v - u| subject to v >= 0 where ``low`` and v <= 0 where ``high``, per axis, by checking
    the two candidates each constrained axis can take (u itself, or the constraint's boundary 0)."""
    best = np.empty_like(u)
    for idx in np.ndindex(u.shape):
        cands = [u[idx], 0.0]
        ok = [c for c in cands if (not low[idx] or c >= 0.0) and (not high[idx] or c <= 0.0)]
        best[idx] = min(ok, key=lambda c: abs(c - u[idx]))
    return best


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_update_grid_is_the_nearest_admissible_velocity(precision):
    res, bc, dt = 10, 3, 1e-3
    rng = np.random.default_rng(4)
    k = O.Kernels(precision)
    R = k.R
    G = res ** 3
    m = rng.uniform(0.0, 2.0, G).astype(R)
    m[rng.random(G) < 0.2] = 0.0
    p = rng.normal(0.0, 1.0, (G, 3)).astype(R)
    g = np.array([0.3, -9.8, 1.7], R)
    v = np.zeros((G, 3), R)
    k.UpdateGrid(O.i64(G), m, p, v, res, bc, g, dt)
    coords = np.stack(np.unravel_index(np.arange(G), (res, res, res)), axis=1)
    low, high = coords < bc, coords >= res - bc
    occ = m > 0
    u = (p[occ] / m[occ, None] + R(dt) * g).astype(R)   # the free velocity at the kernel's precision
    want = _nearest_admissible(u.astype(np.float64), low[occ], high[occ])
    np.testing.assert_array_equal(v[occ].astype(np.float64), want)
    assert np.all(v[~occ] == 0.0)
    # every wall band is exercised both ways: some outward components removed, some inward ones kept
    for band, sign in ((low, -1.0), (high, 1.0)):
        sel = band[occ]
        assert np.any((sign * u[sel] > 0)) and np.any((sign * u[sel] < 0))


def wall_scene(axis, side, speed=1.0):
    """A 0.2 m elastic cube (2x2x2 voxels, 6 tets each, 4 particles per tet) 5 mm from the layers of the wall
    ``side`` (-1 low, +1 high) of ``axis``, moving at ``speed`` into it, no gravity, on the default 40^3 grid."""
    from baking_amd import scenes
    from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
    from baking_amd.geometry.mesh_loader import VoxelTetMesh
    from baking_amd.geometry.mesh_validation import ValidateTetMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    sim = SimulationConfig(gravity=(0.0, 0.0, 0.0), device_name="auto")
    dx, res, bc = sim.grid_spacing, sim.grid_resolution, sim.boundary_cells
    origin = np.full(3, 0.4)
    origin[axis] = bc * dx + 0.005 if side < 0 else (res - bc) * dx - 0.005 - 0.2
    verts, tets = VoxelTetMesh((2, 2, 2), 0.1, origin)
    mesh = ValidateTetMesh(verts, tets)
    v0 = [0.0, 0.0, 0.0]
    v0[axis] = side * speed
    spec = scenes.SceneSpec(f"wall_{'xyz'[axis]}{'-' if side < 0 else '+'}", mesh, SampleParticles(mesh, 1000.0, 4), sim,
                            MaterialConfig(), DingConfig(), TransportConfig(), steps=900, initial_velocity=tuple(v0))
    return spec


def band_violation(node_m, node_v, res, bc, axis, side):
    """Largest outward speed of an occupied node within ``bc`` layers of the wall (0 when none moves outward)."""
    m = np.asarray(node_m).reshape((res,) * 3)
    v = np.asarray(node_v).reshape((res,) * 3 + (3,))
    sl = [slice(None)] * 3
    sl[axis] = slice(0, bc) if side < 0 else slice(res - bc, res)
    mm, vv = m[tuple(sl)], v[tuple(sl)][..., axis]
    occupied = mm > 0
    assert occupied.any(), "the dough reached the wall layers"
    return float(np.max(side * vv[occupied], initial=0.0))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [-1, 1])
def test_cube_stops_at_each_wall(axis, side):
    spec = wall_scene(axis, side)
    o = oracle_from_spec(spec, "f32")
    sim = spec.simulation
    res, bc, dx = sim.grid_resolution, sim.boundary_cells, sim.grid_spacing
    mass = o.m0.astype(np.float64)
    p0 = float(np.dot(mass, o.state.v[:, axis])) * side
    reached = False
    for step in range(1, spec.steps + 1):
        o.Step()
        if step % 50 == 0:
            m = o.node_m.reshape((res,) * 3)
            sl = [slice(None)] * 3
            sl[axis] = slice(0, bc) if side < 0 else slice(res - bc, res)
            if (m[tuple(sl)] > 0).any():
                reached = True
                assert band_violation(o.node_m, o.node_v, res, bc, axis, side) == 0.0, step
    assert reached
    assert not o.flags.any()
    stats = o.check_state()
    assert stats["minimum_j"] > 0
    g = (o.state.x[:, axis].astype(np.float64) - sim.grid_origin[axis]) / dx
    depth = g.min() if side < 0 else (res - 1) - g.max()
    assert depth >= 1.0, depth  # still well inside the valid stencil region (0.5 from the edge)
    p_end = float(np.dot(mass, o.state.v[:, axis])) * side
    assert p_end < -0.1 * p0, (p_end, p0)  # stopped and rebounding from the wall


def sliding_cube(sticky, bottom_cells=0.6):
    """A 0.1 x 0.05 x 0.1 m elastic cube with its bottom at y = bottom_cells dx (inside the floor layers at 0.6), sliding
    along x at 0.5 m/s under gravity on the default 40^3 grid."""
    from baking_amd import scenes
    from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
    from baking_amd.geometry.mesh_loader import VoxelTetMesh
    from baking_amd.geometry.mesh_validation import ValidateTetMesh
    from baking_amd.geometry.particle_sampling import SampleParticles
    sim = SimulationConfig(device_name="auto")
    verts, tets = VoxelTetMesh((4, 2, 4), 0.025, (0.4, bottom_cells * sim.grid_spacing, 0.4))
    mesh = ValidateTetMesh(verts, tets)
    spec = scenes.SceneSpec("sticky_floor" if sticky else "frictionless_floor", mesh, SampleParticles(mesh, 1000.0, 4), sim,
                            MaterialConfig(), DingConfig(), TransportConfig(), steps=50, initial_velocity=(0.5, 0.0, 0.0))
    return spec


@pytest.mark.parametrize("sticky", [True, False])
def test_sticky_floor_closed_form(sticky):
    """Sticky floor (no reference counterpart): every node of the floor's layers has velocity 0, so a particle whose
    whole stencil lies in those layers (grid y < 1.5) keeps its position and deformation gradient exactly; with the
    reference's frictionless floor the same particles keep sliding at about 0.5 m/s."""
    spec = sliding_cube(sticky)
    o = oracle_from_spec(spec, "f32", sticky_floor=sticky)
    sim = spec.simulation
    res, bc, dx = sim.grid_resolution, sim.boundary_cells, sim.grid_spacing
    x0 = o.state.x.copy()
    deep = (x0[:, 1] - sim.grid_origin[1]) / dx < 1.5
    assert deep.sum() >= 8
    for step in range(1, spec.steps + 1):
        o.Step()
        m = o.node_m.reshape((res,) * 3)
        v = o.node_v.reshape((res,) * 3 + (3,))
        floor_nodes = m[:, :bc] > 0
        assert floor_nodes.any()
        if sticky:
            assert np.all(v[:, :bc][floor_nodes] == 0.0), step
            np.testing.assert_array_equal(o.state.x[deep], x0[deep])
            np.testing.assert_array_equal(o.state.F[deep], np.tile(np.eye(3, dtype=np.float32), (int(deep.sum()), 1, 1)))
    if not sticky:
        assert np.all(np.abs(o.state.v[deep, 0] - 0.5) < 0.05)
        assert np.all(o.state.x[deep, 0] > x0[deep, 0] + 0.8 * 0.5 * spec.steps * sim.time_step)
