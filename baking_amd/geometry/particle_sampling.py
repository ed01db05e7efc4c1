"""Particle seeding inside tetrahedra (reference ``geometry/particle_sampling.py:6-53``)."""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class ParticleSamples:
    particle_positions: np.ndarray
    reference_volumes: np.ndarray
    particle_masses: np.ndarray
    source_tets: np.ndarray


# Barycentric weights of the symmetric 4-point tet rule: (5 +- sqrt 5) / 20 layout.
_SYM_OFF = (5.0 - np.sqrt(5.0)) / 20.0
_SYM_ON = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0


def SampleParticles(tet_mesh, density, particles_per_tet=0, target_volume=None, random_seed=42, max_particles=500000):
    """Four symmetric samples per tet, one centroid sample, or seeded Dirichlet samples.

    ``particles_per_tet=0`` picks max(4, ceil(V_tet / target_volume)) samples per tet.
    The result equals the reference's tet-by-tet loop bit for bit, without a Python
    loop over tets: the Dirichlet weights of all randomly sampled tets come from one
    ``default_rng(random_seed).dirichlet`` call (the same stream as the reference's
    per-tet calls, in tet order), and positions are stacked (k x 4) @ (4 x 3)
    products per tet, grouped by sample count, as the reference multiplies them.
    """
    if not np.isfinite(density) or density <= 0.0:
        raise ValueError("density must be positive and finite")
    if particles_per_tet < 0:
        raise ValueError("particles_per_tet must be nonnegative")
    if particles_per_tet == 0:
        if target_volume is None or not np.isfinite(target_volume) or target_volume <= 0.0:
            raise ValueError("automatic sampling needs a positive target_volume")
        raw = np.maximum(4.0, np.ceil(tet_mesh.reference_volumes / target_volume))
        if raw.sum() > max_particles:
            raise ValueError("sampling exceeds max_particles; coarsen the input mesh or sampling")
        counts = raw.astype(np.int64)
    else:
        counts = np.full(len(tet_mesh.tet_indices), particles_per_tet, dtype=np.int64)
    if counts.sum() > max_particles:
        raise ValueError("sampling exceeds max_particles")
    source_tets = np.repeat(np.arange(len(counts)), counts)
    volumes = np.repeat(tet_mesh.reference_volumes / counts, counts)
    positions = np.empty((len(source_tets), 3), dtype=np.float64)
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    corners = np.asarray(tet_mesh.vertex_positions, dtype=np.float64)[tet_mesh.tet_indices]  # (T, 4, 3)
    symmetric = np.full((4, 4), _SYM_OFF)
    np.fill_diagonal(symmetric, _SYM_ON)
    random = (counts != 4) & (counts != 1)
    drawn = None
    if random.any():
        drawn = np.random.default_rng(random_seed).dirichlet(np.ones(4), size=int(counts[random].sum()))
        first_draw = np.concatenate(([0], np.cumsum(counts[random])[:-1]))
        draw_of = np.zeros(len(counts), dtype=np.int64)
        draw_of[random] = first_draw
    for k in np.unique(counts):
        tets = np.flatnonzero(counts == k)
        if k == 4:
            weights = np.broadcast_to(symmetric, (len(tets), 4, 4))
        elif k == 1:
            weights = np.full((len(tets), 1, 4), 0.25)
        else:
            weights = drawn[draw_of[tets][:, None] + np.arange(k)[None, :]]
        rows = starts[tets][:, None] + np.arange(k)[None, :]
        positions[rows.reshape(-1)] = np.matmul(weights, corners[tets]).reshape(-1, 3)
    return ParticleSamples(positions, volumes, density * volumes, source_tets)
