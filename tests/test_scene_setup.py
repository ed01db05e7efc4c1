"""Scene set-up without Python loops over tets (reference geometry/particle_sampling.py:13-53,
physics/thermal_water.py:51-74): the vectorised SampleParticles and BoundaryFaces must equal the oracle's
tet-by-tet restatements of the reference's loops bit for bit, ties and every sampling mode included."""
import numpy as np
import pytest

from baking_amd.geometry.mesh_loader import CreateDemoMesh, CreateSphereMesh
from baking_amd.geometry.mesh_validation import ValidateTetMesh
from baking_amd.geometry.particle_sampling import ParticleSamples, SampleParticles
from baking_amd.physics.thermal_water import BoundaryFaces
from oracle import oracle as O


def _meshes():
    return [("demo", CreateDemoMesh(2)), ("sphere", CreateSphereMesh(0.1, 6))]


@pytest.mark.parametrize("mode", ["symmetric", "centroid", "dirichlet", "automatic"])
@pytest.mark.parametrize("which", [0, 1])
def test_sampling_equals_the_reference_loop(which, mode):
    name, mesh = _meshes()[which]
    ppt, tv = {"symmetric": (4, None), "centroid": (1, None), "dirichlet": (3, None),
               "automatic": (0, float(mesh.reference_volumes.mean()) / 5.5)}[mode]
    got = SampleParticles(mesh, 1000.0, ppt, target_volume=tv, random_seed=7)
    want = O.sample_particles(mesh.vertex_positions, mesh.tet_indices, mesh.reference_volumes, 1000.0, ppt, tv, random_seed=7)
    for a, b in zip((got.particle_positions, got.reference_volumes, got.particle_masses, got.source_tets), want):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("which", [0, 1])
def test_boundary_faces_equal_the_reference_loop(which):
    _, mesh = _meshes()[which]
    for samples in (SampleParticles(mesh, 1000.0, 4), SampleParticles(mesh, 1000.0, 0,
                                                                       target_volume=float(mesh.reference_volumes.mean()) / 7)):
        faces, owners = BoundaryFaces(mesh, samples)
        ref_faces, ref_owners = O.boundary_faces(mesh.vertex_positions, mesh.tet_indices, samples.source_tets,
                                                 samples.particle_positions)
        np.testing.assert_array_equal(faces, ref_faces)
        np.testing.assert_array_equal(owners, ref_owners)
        assert len(faces) > 0


def test_boundary_face_owner_ties_take_the_first_particle():
    """Two particles at the same place are equally close to every face centre: numpy's argmin keeps the first."""
    verts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    mesh = ValidateTetMesh(verts, np.array([[0, 1, 2, 3]]))
    x = np.array([[0.2, 0.2, 0.2], [0.2, 0.2, 0.2], [0.3, 0.1, 0.1]])
    samples = ParticleSamples(x, np.full(3, 1.0 / 18), np.full(3, 1.0), np.zeros(3, dtype=np.int64))
    faces, owners = BoundaryFaces(mesh, samples)
    ref_faces, ref_owners = O.boundary_faces(verts, mesh.tet_indices, samples.source_tets, x)
    np.testing.assert_array_equal(faces, ref_faces)
    np.testing.assert_array_equal(owners, ref_owners)
    assert len(faces) == 4 and 1 not in owners
    # outward: (b - a) x (c - a) . (d - a) <= 0 for the left-out vertex d
    for f in faces:
        d = [v for v in range(4) if v not in f][0]
        a, b, c = verts[f]
        assert np.dot(np.cross(b - a, c - a), verts[d] - a) <= 0


def _inside_own_tet(mesh, positions, source):
    corners = mesh.vertex_positions[mesh.tet_indices[source]]
    lam = np.linalg.solve(np.transpose(corners[:, 1:] - corners[:, :1], (0, 2, 1)), (positions - corners[:, 0])[..., None])[..., 0]
    return np.all(lam >= -1e-12, axis=1) & (lam.sum(axis=1) <= 1.0 + 1e-12)


def test_oracle_sampling_pins():
    """oracle.sample_particles against the reference's own checks (tests/test_mesh_processing.py:8-21, 53-59):
    CreateDemoMesh(2) volume 0.3^3, mass 27 at density 1000, every sample inside its tet; automatic sampling of
    CreateDemoMesh(1) at target volume 0.001 gives at least 24 particles of total volume 0.027."""
    mesh = CreateDemoMesh(2)
    np.testing.assert_allclose(mesh.reference_volumes.sum(), 0.3 ** 3)
    for ppt in (4, 1, 3):
        x, vol, mass, src = O.sample_particles(mesh.vertex_positions, mesh.tet_indices, mesh.reference_volumes, 1000.0, ppt)
        np.testing.assert_allclose(vol.sum(), 0.3 ** 3)
        np.testing.assert_allclose(mass.sum(), 27.0)
        assert np.all(_inside_own_tet(mesh, x, src))
        assert np.all(np.bincount(src, minlength=len(mesh.tet_indices)) == ppt)
    small = CreateDemoMesh(1)
    x, vol, mass, src = O.sample_particles(small.vertex_positions, small.tet_indices, small.reference_volumes, 1000.0, 0, 0.001)
    assert len(x) >= 24
    np.testing.assert_allclose(vol.sum(), 0.027)
    assert np.all(_inside_own_tet(small, x, src))


def _quadratic(x, c):
    return (c[0] + x @ c[1:4] + (x[:, 0] ** 2) * c[4] + (x[:, 1] ** 2) * c[5] + (x[:, 2] ** 2) * c[6]
            + x[:, 0] * x[:, 1] * c[7] + x[:, 1] * x[:, 2] * c[8] + x[:, 0] * x[:, 2] * c[9])


def test_oracle_sampling_rules_integrate_exactly():
    """The symmetric 4-point rule integrates every quadratic exactly over each tet, the centroid rule every linear
    function.  The exact integral of a quadratic f over a tet of volume V is V (-sum_i f(v_i) / 20
    + sum_edges f(m_ij) / 5), v_i its vertices and m_ij its edge midpoints (the quadratic Lagrange element's
    weights), so sum_p V_p f(x_p) must equal the sum of these over the mesh."""
    mesh = CreateSphereMesh(0.1, 3)
    rng = np.random.default_rng(9)
    c = rng.normal(size=10)
    corners = mesh.vertex_positions[mesh.tet_indices]
    V = mesh.reference_volumes
    vert = sum(_quadratic(corners[:, i], c) for i in range(4))
    edge = sum(_quadratic(0.5 * (corners[:, i] + corners[:, j]), c) for i in range(4) for j in range(i + 1, 4))
    exact = float(np.sum(V * (-vert / 20.0 + edge / 5.0)))
    x, vol, _, _ = O.sample_particles(mesh.vertex_positions, mesh.tet_indices, V, 1000.0, 4)
    np.testing.assert_allclose(np.dot(vol, _quadratic(x, c)), exact, rtol=1e-12)
    lin = c.copy()
    lin[4:] = 0.0
    x1, vol1, _, _ = O.sample_particles(mesh.vertex_positions, mesh.tet_indices, V, 1000.0, 1)
    np.testing.assert_allclose(np.dot(vol1, _quadratic(x1, lin)), float(np.sum(V * _quadratic(corners.mean(axis=1), lin))),
                               rtol=1e-12)
    # the centroid rule is not exact for a quadratic: the test above would catch a swapped rule
    assert abs(np.dot(vol1, _quadratic(x1, c)) - exact) > 1e-9 * abs(exact)
