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
