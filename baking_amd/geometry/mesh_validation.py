"""Tetrahedral mesh container and validation (reference ``geometry/mesh_validation.py:1-47``).

``ValidateTetMesh`` runs a fixed sequence of stages over the input -- array shapes, index type and range,
vertex finiteness, repeated and duplicate tetrahedra, degeneracy -- each raising the reference's message,
then orients every tet positively by exchanging its first two vertices.  Reference volumes are
|det([v1 - v0, v2 - v0, v3 - v0])| / 6, the reference's definition (the particles' volumes are split from
them, so they keep its rounding).
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class TetMesh:
    vertex_positions: np.ndarray
    tet_indices: np.ndarray
    reference_volumes: np.ndarray


def ComputeSignedVolumes(vertex_positions, tet_indices):
    """Signed tet volumes det([v1 - v0, v2 - v0, v3 - v0]) / 6 (positive when v1, v2, v3 wind counter-clockwise seen
    from outside opposite v0)."""
    anchor = vertex_positions[tet_indices[:, 0]]
    edges = np.stack([vertex_positions[tet_indices[:, k]] - anchor for k in (1, 2, 3)], axis=1)
    return np.linalg.det(edges) / 6.0


def _longest_edge(vertex_positions, tet_indices):
    """Per tet, the longest of its six edges."""
    pairs = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
    lengths = [np.linalg.norm(vertex_positions[tet_indices[:, a]] - vertex_positions[tet_indices[:, b]], axis=1)
               for a, b in pairs]
    return np.max(lengths, axis=0)


def _stage_shapes(vertices, raw):
    if vertices.ndim != 2 or vertices.shape[1] != 3:
        return "vertices must have shape (vertex_count, 3)"
    if raw.ndim != 2 or raw.shape[1] != 4 or raw.shape[0] == 0:
        return "tetrahedra must have nonempty shape (tet_count, 4)"
    if raw.dtype.kind not in "iu":
        return "tetrahedron indices must be integers (zero based)"
    if vertices.shape[0] < 4 or not np.all(np.isfinite(vertices)):
        return "vertices must be finite and contain at least four entries"
    if raw.min() < 0 or raw.max() >= vertices.shape[0]:
        return "tetrahedron index outside vertex array"
    return None


def _stage_topology(tets):
    canonical = np.sort(tets, axis=1)
    if np.any(canonical[:, 1:] == canonical[:, :-1]):
        return "tetrahedron contains repeated vertex indices"
    if np.unique(canonical, axis=0).shape[0] != tets.shape[0]:
        return "duplicate tetrahedra detected"
    return None


def ValidateTetMesh(vertex_positions, tet_indices):
    """A :class:`TetMesh` of the input with every tet positively oriented, or ValueError naming the first problem."""
    vertices = np.array(vertex_positions, dtype=np.float64, order="C")
    raw = np.asarray(tet_indices)
    problem = _stage_shapes(vertices, raw)
    if problem is None:
        tets = raw.astype(np.int32, order="C")
        problem = _stage_topology(tets)
    if problem is not None:
        raise ValueError(problem)
    signed = ComputeSignedVolumes(vertices, tets)
    if np.any(np.abs(signed) <= 1.0e-12 * _longest_edge(vertices, tets) ** 3):
        raise ValueError("zero-volume or nearly degenerate tetrahedron detected")
    inverted = signed < 0.0
    tets[inverted] = tets[inverted][:, [1, 0, 2, 3]]  # same vertices, opposite winding
    return TetMesh(vertices, tets, np.abs(signed))
