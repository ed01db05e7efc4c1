"""Tet mesh input and built-in meshes (reference ``geometry/mesh_loader.py:6-79``)."""
from pathlib import Path

import numpy as np

from .mesh_validation import ValidateTetMesh

# Six tetrahedra per voxel around the main diagonal 0-7 (corner bit order x, y, z).
_VOXEL_TETS = np.array([[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7], [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]])
_VOXEL_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]])


def LoadTetMesh(mesh_path, mesh_scale=1.0, mesh_translation=(0.0, 0.0, 0.0)):
    """Read an NPZ with ``vertices``/``tetrahedra`` (or a meshio format when meshio exists)."""
    mesh_path = Path(mesh_path)
    if not np.isfinite(mesh_scale) or mesh_scale <= 0.0:
        raise ValueError("mesh_scale must be positive and finite")
    translation = np.asarray(mesh_translation, dtype=np.float64)
    if translation.shape != (3,) or not np.isfinite(translation).all():
        raise ValueError("mesh_translation must contain three finite values")
    if mesh_path.suffix.lower() == ".npz":
        with np.load(mesh_path, allow_pickle=False) as data:
            vertices = data["vertices"]
            tets = data["tetrahedra"]
    else:
        import meshio  # optional; absent on this image
        mesh = meshio.read(mesh_path)
        if any(b.type.startswith("tetra") and b.type != "tetra" for b in mesh.cells):
            raise ValueError("higher-order tetrahedra are unsupported; export linear tetrahedra")
        blocks = [b.data for b in mesh.cells if b.type == "tetra"]
        if not blocks:
            raise ValueError("input contains no linear tetrahedra")
        tets = np.concatenate(blocks)
        vertices = mesh.points
    vertices = np.asarray(vertices, dtype=np.float64) * mesh_scale + translation
    return ValidateTetMesh(vertices, tets)


def VoxelTetMesh(shape, spacing, origin):
    """Structured box of ``shape`` voxels, six tets per voxel; vertex index is row-major (x, y, z)."""
    shape = tuple(int(s) for s in shape)
    axes = [np.arange(n + 1) * spacing + o for n, o in zip(shape, origin)]
    vertices = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    node_shape = tuple(n + 1 for n in shape)
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), axis=-1).reshape(-1, 3)
    corner_ids = np.stack([np.ravel_multi_index((cells + c).T, node_shape) for c in _VOXEL_CORNERS], axis=1)
    tets = corner_ids[:, _VOXEL_TETS].reshape(-1, 4)
    return vertices, tets


def CreateDemoMesh(subdivision_count=6):
    """A 0.3 m cube at (0.35, 0.45, 0.35), six tets per voxel."""
    if subdivision_count < 1:
        raise ValueError("subdivision_count must be positive")
    n = subdivision_count
    vertices, tets = VoxelTetMesh((n, n, n), 0.3 / n, (0.0, 0.0, 0.0))
    vertices = np.linspace(0.0, 0.3, n + 1)[np.indices((n + 1,) * 3).reshape(3, -1).T]
    vertices = vertices + np.array([0.35, 0.45, 0.35])
    return ValidateTetMesh(vertices, tets)


def CreateSphereMesh(radius=0.1, subdivision_count=12):
    """Map the structured cube onto a ball of ``radius`` centred at the origin.

    Uses the cube-to-sphere map x*sqrt(1 - y^2/2 - z^2/2 + y^2 z^2/3) per axis.
    """
    cube = CreateDemoMesh(subdivision_count)
    low = cube.vertex_positions.min(axis=0)
    extent = np.ptp(cube.vertex_positions, axis=0)
    unit = 2.0 * (cube.vertex_positions - low) / extent - 1.0
    ball = np.empty_like(unit)
    for axis in range(3):
        a, b = [k for k in range(3) if k != axis]
        ya, zb = unit[:, a] ** 2, unit[:, b] ** 2
        ball[:, axis] = radius * unit[:, axis] * np.sqrt(1.0 - ya / 2.0 - zb / 2.0 + ya * zb / 3.0)
    return ValidateTetMesh(ball, cube.tet_indices)
