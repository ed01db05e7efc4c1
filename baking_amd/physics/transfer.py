"""MLS-MPM particle/grid transfers (reference ``physics/transfer.py:1-109``).

``ParticleToGrid`` and ``GridToParticle`` run the gfx950 kernels on the
reference's dense node layout; the fused substep (``bk_substep``) uses the same
arithmetic on a tiled grid.  The stencil helpers are host restatements for
scene set-up and checks.
"""
import math

import numpy as np

from ..launch import Kernel

ParticleToGrid = Kernel("bk_ParticleToGrid", ["f32"] * 8 + ["v3", "f", "i", "f", "i32"])
GridToParticle = Kernel("bk_GridToParticle", ["f32"] * 7 + ["v3", "f", "i", "f", "i32"])


def quadratic_weight(local_coord, stencil_index):
    """Quadratic B-spline weight of stencil node 0, 1 or 2 (transfer.py:4-13)."""
    if stencil_index == 0:
        return 0.5 * (1.5 - local_coord) ** 2
    if stencil_index == 1:
        return 0.75 - (local_coord - 1.0) ** 2
    return 0.5 * (local_coord - 0.5) ** 2


def stencil_base(grid_position):
    return tuple(int(math.floor(g - 0.5)) for g in grid_position)


def valid_stencil(base_coord, grid_resolution):
    return all(b >= 0 for b in base_coord) and all(b + 2 < grid_resolution for b in base_coord)


def flatten_node(node_coord, grid_resolution):
    i, j, k = node_coord
    return (i * grid_resolution + j) * grid_resolution + k


def cell_keys_host(positions, origin, spacing, resolution):
    """Row-major cell key flatten_node(stencil_base((x - o) / dx)) in float32 arithmetic."""
    g = (np.asarray(positions, np.float32) - np.asarray(origin, np.float32)) / np.float32(spacing)
    base = np.floor(g - np.float32(0.5)).astype(np.int64)
    return (base[:, 0] * resolution + base[:, 1]) * resolution + base[:, 2]
