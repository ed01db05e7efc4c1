"""Tetrahedral mesh container and validation (reference ``geometry/mesh_validation.py:1-47``)."""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class TetMesh:
    vertex_positions: np.ndarray
    tet_indices: np.ndarray
    reference_volumes: np.ndarray


def ComputeSignedVolumes(vertex_positions, tet_indices):
    """Signed tet volumes det([v1-v0, v2-v0, v3-v0]) / 6."""
    corners = vertex_positions[tet_indices]
    edges = corners[:, 1:] - corners[:, :1]
    return np.linalg.det(edges) / 6.0


def ValidateTetMesh(vertex_positions, tet_indices):
    """Check shapes, indices, duplicates and degeneracy; orient every tet positively."""
    vertices = np.asarray(vertex_positions, dtype=np.float64)
    raw = np.asarray(tet_indices)
    if vertices.ndim != 2 or vertices.shape[1] != 3:
        raise ValueError("vertices must have shape (vertex_count, 3)")
    if raw.ndim != 2 or raw.shape[1] != 4 or len(raw) == 0:
        raise ValueError("tetrahedra must have nonempty shape (tet_count, 4)")
    if not np.issubdtype(raw.dtype, np.integer):
        raise ValueError("tetrahedron indices must be integers (zero based)")
    if len(vertices) < 4 or not np.isfinite(vertices).all():
        raise ValueError("vertices must be finite and contain at least four entries")
    if raw.min() < 0 or raw.max() >= len(vertices):
        raise ValueError("tetrahedron index outside vertex array")
    tets = raw.astype(np.int32, copy=True)
    ordered = np.sort(tets, axis=1)
    if np.any(np.diff(ordered, axis=1) == 0):
        raise ValueError("tetrahedron contains repeated vertex indices")
    if len(np.unique(ordered, axis=0)) != len(tets):
        raise ValueError("duplicate tetrahedra detected")
    volumes = ComputeSignedVolumes(vertices, tets)
    corners = vertices[tets]
    edge_scale = np.max(np.linalg.norm(corners[:, :, None] - corners[:, None, :], axis=-1), axis=(1, 2))
    if np.any(np.abs(volumes) <= 1.0e-12 * edge_scale ** 3):
        raise ValueError("zero-volume or nearly degenerate tetrahedron detected")
    flip = volumes < 0.0
    tets[flip, 0], tets[flip, 1] = tets[flip, 1].copy(), tets[flip, 0].copy()
    return TetMesh(vertices.copy(), tets, np.abs(volumes))
