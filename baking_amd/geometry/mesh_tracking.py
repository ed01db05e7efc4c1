"""Tet-mesh reconstruction from particle positions and total F (reference ``geometry/mesh_tracking.py:1-81``).

Each vertex is bound to the particles sampled in its incident tets; a particle
predicts x_v = x_p + F_p (X_v - X_p) and the predictions are blended with
normalised reference-volume weights, summed as deviations from an anchor.
"""
import numpy as np
import torch

from ..launch import Kernel

ReconstructMeshVertices = Kernel("bk_ReconstructMeshVertices", ["f32", "f32", "f32", "i32", "i32", "f32", "f32", "f32"])


def BuildBinding(tet_mesh, particle_samples):
    """Host CSR vertex -> particles with weights and reference offsets (mesh_tracking.py:42-75)."""
    nv = len(tet_mesh.vertex_positions)
    n = len(particle_samples.particle_positions)
    source_tets = np.asarray(particle_samples.source_tets)
    volumes = np.asarray(particle_samples.reference_volumes, dtype=np.float64)
    if source_tets.shape != (n,) or not np.issubdtype(source_tets.dtype, np.integer):
        raise ValueError("source_tets must contain one integer tet index per particle")
    if n == 0 or source_tets.min() < 0 or source_tets.max() >= len(tet_mesh.tet_indices):
        raise ValueError("invalid particle source tetrahedra")
    if volumes.shape != (n,) or not np.isfinite(volumes).all() or np.any(volumes <= 0.0):
        raise ValueError("particle reference volumes must be positive and finite")
    if np.any(np.bincount(source_tets, minlength=len(tet_mesh.tet_indices)) == 0):
        raise ValueError("every tetrahedron must contain particles for mesh reconstruction")
    vertices = tet_mesh.tet_indices[source_tets].reshape(-1)
    particles = np.repeat(np.arange(n, dtype=np.int32), 4)
    order = np.argsort(vertices, kind="stable")
    vertices, particles = vertices[order], particles[order]
    counts = np.bincount(vertices, minlength=nv)
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    weights = volumes[particles]
    weights = weights / np.bincount(vertices, weights=weights, minlength=nv)[vertices]
    offsets_xyz = tet_mesh.vertex_positions[vertices] - particle_samples.particle_positions[particles]
    return dict(offsets=offsets, particles=particles.astype(np.int32), weights=weights.astype(np.float32),
                reference_offsets=offsets_xyz.astype(np.float32), unmapped=int(np.count_nonzero(counts == 0)))


class ParticleMeshBinding:
    def __init__(self, tet_mesh, particle_samples, device_name):
        b = BuildBinding(tet_mesh, particle_samples)
        self.unmapped_vertex_count = b["unmapped"]
        dev = device_name
        self.reference_vertices = torch.as_tensor(tet_mesh.vertex_positions.astype(np.float32), device=dev)
        self.binding_offsets = torch.as_tensor(b["offsets"], device=dev)
        self.binding_particles = torch.as_tensor(b["particles"], device=dev)
        self.binding_weights = torch.as_tensor(b["weights"], device=dev)
        self.reference_offsets = torch.as_tensor(b["reference_offsets"], device=dev)

    def Reconstruct(self, particle_state, output_vertices):
        ReconstructMeshVertices(len(output_vertices), particle_state.particle_positions, particle_state.deformation_gradients,
                                self.reference_vertices, self.binding_offsets, self.binding_particles, self.binding_weights,
                                self.reference_offsets, output_vertices)
