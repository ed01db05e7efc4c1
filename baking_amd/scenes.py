"""Synthetic dough scenes: one jittered-seeded dough plus the tet mesh its transport needs.

Workloads (all SI units):

* ``sphere_small``   -- sphere r = 0.1 at (0.5, 0.5, 0.3), 64^3 grid (dx = 1/64), 8 particles per cell
  (2x2x2 jittered, seed 0) inside the sphere (~8.8k particles), dt = 1e-4, oven 453.15 K.
* ``loaf_256``       -- box 0.30 x 0.15 x 0.12 on a 256^3 grid (dx = 1/256), 8 particles per cell (~0.72M),
  moisture 0.4 + 0.05 * Gaussian random field (correlation length 0.02, seed 1), porosity U[0.03, 0.07],
  oven 473.15 K, dt = 5e-5.
* ``batched_doughs`` -- 48 independent spheres, boxes and ellipsoids (half-extent 0.05-0.15, seed 2) on their own
  64^3 or 128^3 grids, 3k-60k particles each (~1.0M total), material as ``loaf_256``, dt = 1.25e-5.

Parameters the workload text leaves open are taken from the reference's
example scene (``examples/oven_baking.txt``) and recorded in ``SceneSpec``.
Seeding: x = dx (c + (s + 1/2 + u) / 2), s in {0,1}^3, u ~ U(-1/2, 1/2)^3 drawn
cell by cell in row-major cell order from ``default_rng(seed)``; particles
outside the body are dropped.  Every particle is assigned to the tet that
contains it (or the nearest one for particles just outside the faceted
surface), and tets left empty take the nearest particle from a tet that has
several, so mesh reconstruction has a binding for every tet.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
from .geometry.mesh_loader import CreateSphereMesh, VoxelTetMesh
from .geometry.mesh_validation import ValidateTetMesh
from .geometry.particle_sampling import ParticleSamples
from .physics.materials.gas import gas_constant


@dataclass
class SceneSpec:
    name: str
    tet_mesh: object
    samples: ParticleSamples
    simulation: SimulationConfig
    material: MaterialConfig
    ding: DingConfig
    transport: TransportConfig
    steps: int
    initial_velocity: tuple = (0.0, 0.0, 0.0)
    # per-particle initial fields assigned after construction (None = config defaults)
    water_masses: np.ndarray | None = None
    porosities: np.ndarray | None = None
    gas_moles: np.ndarray | None = None
    velocities: np.ndarray | None = None
    affine_velocities: np.ndarray | None = None
    temperatures: np.ndarray | None = None
    notes: dict = field(default_factory=dict)


def jittered_cells(inside, lo, hi, dx, seed):
    """8 jittered particles per cell (2x2x2) for cells overlapping [lo, hi]; keep those with inside(x)."""
    c0 = np.floor(np.asarray(lo) / dx).astype(int)
    c1 = np.ceil(np.asarray(hi) / dx).astype(int)
    shape = c1 - c0
    cells = np.stack(np.meshgrid(*[np.arange(a, b) for a, b in zip(c0, c1)], indexing="ij"), axis=-1).reshape(-1, 3)
    sub = np.stack(np.meshgrid([0, 1], [0, 1], [0, 1], indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.5, 0.5, size=(len(cells), 8, 3))
    pos = dx * (cells[:, None, :] + (sub[None, :, :] + 0.5 + u) / 2.0)
    pos = pos.reshape(-1, 3)
    keep = inside(pos)
    del shape
    return pos[keep]


def assign_tets(tet_mesh, positions, k=12):
    """Containing tet for every point (nearest by barycentric margin when outside), then fill empty tets."""
    from scipy.spatial import cKDTree
    verts = tet_mesh.vertex_positions
    tets = tet_mesh.tet_indices
    corners = verts[tets]
    origin = corners[:, 0]
    edges = np.transpose(corners[:, 1:] - origin[:, None, :], (0, 2, 1))  # (T,3,3) columns = edges
    inv = np.linalg.inv(edges)
    centroids = corners.mean(axis=1)
    tree = cKDTree(centroids)
    k = min(k, len(tets))
    _, cand = tree.query(positions, k=k)
    cand = cand.reshape(len(positions), k)
    rel = positions[:, None, :] - origin[cand]
    lam = np.einsum("nkij,nkj->nki", inv[cand], rel)
    bary = np.concatenate([1.0 - lam.sum(axis=2, keepdims=True), lam], axis=2)
    margin = bary.min(axis=2)
    best = cand[np.arange(len(positions)), np.argmax(margin, axis=1)]
    source = best.astype(np.int64)
    counts = np.bincount(source, minlength=len(tets))
    empty = np.flatnonzero(counts == 0)
    if len(empty):
        ptree = cKDTree(positions)
        for t in empty:
            _, near = ptree.query(centroids[t], k=min(64, len(positions)))
            for p in np.atleast_1d(near):
                if counts[source[p]] >= 2:
                    counts[source[p]] -= 1
                    source[p] = t
                    counts[t] += 1
                    break
            else:
                raise ValueError("could not give every tet a particle; refine seeding or coarsen the mesh")
    return source


def make_samples(tet_mesh, positions, dx, density):
    """Particles of volume dx^3 / 8 and mass density * volume, tagged with their source tet."""
    n = len(positions)
    volumes = np.full(n, dx ** 3 / 8.0)
    source = assign_tets(tet_mesh, positions)
    return ParticleSamples(positions, volumes, density * volumes, source)


# Ding yield stress of the scaled workloads (the loaf and the batched doughs): twice the example's 810.82 Pa.  At
# the example's value the loaf's weight (rho g h ~ 1.7 kPa) compacts its base until the pore gas stiffens past
# the gas-acoustic CFL bound at substep ~806 of dt = 5e-5, whatever the gas production or the modulus; at 1600 Pa
# the base still yields (porosity down to ~0.017) and all 2000 substeps stay below CFL 0.31.
SCALED_YIELD_STRESS = 1600.0


def _oven_material(young=1.25e5, poisson=0.485):
    return MaterialConfig(density=1150.0, young_modulus=young, poisson_ratio=poisson, model="ding2019")


def _oven_ding(porosity, p0, gas_production=2.04e-5, yield_stress=810.82, viscosity=1.5e6):
    return DingConfig(yield_stress=yield_stress, viscosity=viscosity, porosity=porosity, temperature_enabled=True,
                      temperature_knots_k=(278.15, 300.15, 328.15, 353.15), yield_factors=(1.5, 1.0, 0.5, 3.0),
                      viscosity_factors=(1.5, 1.0, 0.5, 5.0), gas_enabled=True, gas_initial_pressure_pa=p0,
                      gas_ambient_pressure_pa=p0, gas_production_rate=gas_production, gas_reaction_temperature_k=318.15,
                      gas_reaction_width_k=15.42)


def _oven_transport(T0, T_oven, s0=0.4):
    return TransportConfig(enabled=True, water_enabled=True, initial_temperature_k=T0, oven_temperature_k=T_oven,
                           conductivity=0.36, specific_heat=2830.0, heat_transfer=35.0, tray_temperature_k=T_oven,
                           tray_heat_transfer=100.0, contact_distance=0.001, initial_saturation=s0, water_density=1000.0,
                           water_diffusivity=1.5e-10, surface_water_loss=5.0e-7, water_stress=10000.0, latent_heat=2.26e6)


def sphere_small(seed=0, mesh_subdivisions=6):
    """The small plain scene: sphere r = 0.1 at (0.5, 0.5, 0.3), 64^3 grid, ~8.8k particles."""
    res, dx = 64, 1.0 / 64.0
    centre, radius = np.array([0.5, 0.5, 0.3]), 0.1
    sphere = CreateSphereMesh(radius, mesh_subdivisions)
    mesh = ValidateTetMesh(sphere.vertex_positions + centre, sphere.tet_indices)
    pos = jittered_cells(lambda p: np.linalg.norm(p - centre, axis=1) <= radius, centre - radius, centre + radius, dx, seed)
    material = _oven_material()
    samples = make_samples(mesh, pos, dx, material.density)
    sim = SimulationConfig(grid_resolution=res, grid_spacing=dx, grid_origin=(0.0, 0.0, 0.0), time_step=1.0e-4,
                           gravity=(0.0, -9.8, 0.0), boundary_cells=3, device_name="auto")
    T0 = 298.15
    return SceneSpec("sphere_small", mesh, samples, sim, material, _oven_ding(0.05, 101325.0),
                     _oven_transport(T0, 453.15), steps=1000,
                     notes=dict(seed=seed, mesh=f"CreateSphereMesh(0.1, {mesh_subdivisions}) + (0.5, 0.5, 0.3)",
                                parameters="examples/oven_baking.txt material/ding/transport; n0 = 0.05, T0 = 298.15 K"))


def sphere_gradient(target_seed=3):
    """The differentiable workload: the small sphere for 256 substeps, checkpointed every 16.

    Loss L = (y_c - y*)^2 + (s_mean - s*)^2 of the final state, with y_c the mass-weighted centroid height
    (mass = dry mass + water) and s_mean the mean saturation W / (rho_w V0 det F); the targets are the
    initial centroid height and saturation moved by U(-0.01, 0.01) each from ``default_rng(target_seed)``.
    The parameters are each particle's initial porosity theta (its gas content following as
    p0 theta / (R T0)) and the oven temperature.  ``notes`` holds the targets and sizes."""
    spec = sphere_small()
    spec.name = "sphere_gradient"
    spec.steps = 256
    rng = np.random.default_rng(target_seed)
    x = spec.samples.particle_positions
    t = spec.transport
    V0 = np.asarray(spec.samples.reference_volumes, np.float64)
    dry = (spec.material.density - t.water_density * t.initial_saturation) * V0
    m = dry + t.water_density * t.initial_saturation * V0
    y0 = float(np.dot(m, x[:, 1]) / m.sum())
    spec.notes.update(checkpoint_every=16, centroid_target=y0 + float(rng.uniform(-0.01, 0.01)),
                      saturation_target=float(t.initial_saturation) + float(rng.uniform(-0.01, 0.01)))
    return spec


def gradient_loss(spec, positions, water, gradients, xp=np):
    """L of ``sphere_gradient`` from final positions (N x 3), water masses (N) and deformation gradients
    (N x 3 x 3); ``xp`` is numpy or torch (then differentiable; float64 inside)."""
    t = spec.transport
    V0 = np.asarray(spec.samples.reference_volumes, np.float64)
    dry_np = (spec.material.density - t.water_density * t.initial_saturation) * V0
    if xp is np:
        W = np.asarray(water, np.float64)
        y = np.asarray(positions, np.float64)[:, 1]
        J = np.linalg.det(np.asarray(gradients, np.float64))
        dry, vol = dry_np, V0
    else:
        W = water.double()
        y = positions[:, 1].double()
        J = xp.linalg.det(gradients.double())
        dry = xp.as_tensor(dry_np, device=W.device)
        vol = xp.as_tensor(V0, device=W.device)
    m = dry + W
    yc = (m * y).sum() / m.sum()
    sat = (W / (float(t.water_density) * vol * J)).mean()
    return (yc - spec.notes["centroid_target"]) ** 2 + (sat - spec.notes["saturation_target"]) ** 2


def gaussian_random_field(shape, spacing, correlation_length, rng):
    """Unit-variance periodic Gaussian random field: white noise filtered by exp(-k^2 l^2 / 4) in Fourier space."""
    noise = rng.standard_normal(shape)
    spec = np.fft.rfftn(noise)
    ks = [np.fft.fftfreq(n, d=spacing) * 2 * np.pi for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1], d=spacing) * 2 * np.pi]
    k2 = sum(np.meshgrid(*[k ** 2 for k in ks], indexing="ij"))
    field = np.fft.irfftn(spec * np.exp(-k2 * correlation_length ** 2 / 4.0), s=shape)
    field -= field.mean()
    return field / field.std()


def trilinear(field, origin, spacing, points):
    g = (points - origin) / spacing
    i0 = np.clip(np.floor(g).astype(int), 0, np.array(field.shape) - 2)
    f = g - i0
    out = np.zeros(len(points))
    for dx_ in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx_ else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out += w * field[i0[:, 0] + dx_, i0[:, 1] + dy, i0[:, 2] + dz]
    return out


def loaf_256(seed=0, field_seed=1, voxel=0.015, p0=1000.0, young=1.25e5, gas_production=2.04e-5, steps=2000,
             yield_stress=SCALED_YIELD_STRESS, viscosity=1.5e6):
    """The scaled plain scene: 0.30 x 0.15 x 0.12 loaf on a 256^3 grid, ~0.72M particles.

    The pore-gas pressure p0 = ambient = 1000 Pa keeps the gas-acoustic CFL
    below 0.4 at dx = 1/256, dt = 5e-5 (atmospheric pore gas would trip it on
    the first substep); Poisson's ratio 0.45 keeps the elastic CFL below 0.4;
    the yield stress SCALED_YIELD_STRESS keeps the base's compaction, and with
    it the pore gas's stiffness, below the gas-acoustic bound for all 2000
    substeps (the example's 810.82 Pa trips it at substep ~806).
    """
    res, dx = 256, 1.0 / 256.0
    size = np.array([0.30, 0.15, 0.12])
    floor = (3 - 1) * dx
    lo = np.array([0.5 - size[0] / 2, floor + 1.0e-4, 0.5 - size[2] / 2])
    hi = lo + size
    pos = jittered_cells(lambda p: np.all((p >= lo) & (p <= hi), axis=1), lo, hi, dx, seed)
    shape = np.round(size / voxel).astype(int)
    verts, tets = VoxelTetMesh(shape, voxel, lo)
    verts = lo + (verts - lo) * (size / (shape * voxel))
    mesh = ValidateTetMesh(verts, tets)
    material = _oven_material(young=young, poisson=0.45)
    samples = make_samples(mesh, pos, dx, material.density)
    rng = np.random.default_rng(field_seed)
    margin = 0.06
    g_lo = lo - margin
    g_shape = tuple(int(v) for v in np.ceil((size + 2 * margin) / dx) + 2)
    grf = gaussian_random_field(g_shape, dx, 0.02, rng)
    s = 0.4 + 0.05 * trilinear(grf, g_lo, dx, pos)
    porosity = rng.uniform(0.03, 0.07, size=len(pos))
    T0 = 298.15
    water = 1000.0 * s * samples.reference_volumes
    moles = p0 * porosity / (gas_constant * T0)
    sim = SimulationConfig(grid_resolution=res, grid_spacing=dx, grid_origin=(0.0, 0.0, 0.0), time_step=5.0e-5,
                           gravity=(0.0, -9.8, 0.0), boundary_cells=3, device_name="auto")
    return SceneSpec("loaf_256", mesh, samples, sim, material, _oven_ding(0.05, p0, gas_production, yield_stress, viscosity),
                     _oven_transport(T0, 473.15),
                     steps=steps, water_masses=water, porosities=porosity.astype(np.float32), gas_moles=moles,
                     notes=dict(seed=seed, field_seed=field_seed, p0=p0, poisson=0.45, voxel=voxel, young=young,
                                gas_production=gas_production, yield_stress=yield_stress, viscosity=viscosity))


def dough_shapes(count=48, seed=2, half_extent=(0.05, 0.15), max_particles=60000):
    """Shapes, half-extents and grids of a batch of uneven doughs, drawn from ``default_rng(seed)``.

    Each dough is a sphere, box or ellipsoid (uniform choice) with half-extent
    h ~ U[half_extent]: radius h, or half-edges / semi-axes h * U[0.6, 1.0] per
    axis.  It gets a 128^3 grid when 8 particles per cell of it stay within
    ``max_particles``, else a 64^3 grid.  Returns a list of dicts.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        shape = ("sphere", "box", "ellipsoid")[int(rng.integers(3))]
        h = float(rng.uniform(*half_extent))
        axes = np.full(3, h) if shape == "sphere" else h * rng.uniform(0.6, 1.0, 3)
        volume = 8.0 * np.prod(axes) if shape == "box" else 4.0 / 3.0 * np.pi * np.prod(axes)
        res = 128 if 8.0 * volume * 128 ** 3 <= max_particles else 64
        out.append(dict(shape=shape, half_extent=h, axes=axes, res=res, expected_particles=int(8.0 * volume * res ** 3)))
    return out


def dough_scene(shape, axes, res, seed, name="dough", p0=1000.0, porosity_range=(0.03, 0.07), dt=1.25e-5):
    """One dough resting on the tray of a unit-cube oven: sphere, box or ellipsoid with semi-axes ``axes``.

    Material as ``loaf_256`` (Poisson's ratio 0.45, pore gas at ambient p0 =
    1000 Pa, oven 473.15 K), moisture at the configured initial saturation,
    porosity U[porosity_range] per particle; particles jittered 8 per cell from
    ``default_rng(seed)``, porosities from the same generator's next draws.
    dt = 1.25e-5: a rounded dough on the tray compacts at its contact patch
    (porosity down to ~1e-3 after 0.02 s), which stiffens its pore gas until
    the gas-acoustic CFL check (gas.py:34-35) fails -- from t = 0.02 s at
    dt = 5e-5 and within 1000 substeps at 2.5e-5; 1000 substeps at 1.25e-5
    (12.5 ms of baking) stay clear of it.
    """
    dx = 1.0 / res
    axes = np.asarray(axes, dtype=np.float64)
    floor = (3 - 1) * dx
    centre = np.array([0.5, floor + 1.0e-4 + axes[1], 0.5])
    lo, hi = centre - axes, centre + axes
    # mesh resolution: at most about one tet per 8 expected particles, so every tet gets one
    if shape == "box":
        inside = lambda p: np.all(np.abs(p - centre) <= axes, axis=1)  # noqa: E731
        size = 2.0 * axes
        expected = 8.0 * np.prod(size) / dx ** 3
        per_axis = int(np.clip(np.floor((expected / 48.0) ** (1.0 / 3.0)), 1, 6))
        counts = np.maximum(1, np.round(size / (size.max() / per_axis))).astype(int)
        verts, tets = VoxelTetMesh(counts, 1.0, (0.0, 0.0, 0.0))
        verts = lo + verts * (size / counts)
    else:
        inside = lambda p: np.sum(((p - centre) / axes) ** 2, axis=1) <= 1.0  # noqa: E731
        expected = 8.0 * (4.0 / 3.0) * np.pi * np.prod(axes) / dx ** 3
        tets_by_subdivision = {2: 48, 3: 162, 4: 384, 5: 750}
        subdivision = max([k for k, t in tets_by_subdivision.items() if 8 * t <= expected] or [2])
        unit = CreateSphereMesh(1.0, subdivision)
        verts, tets = unit.vertex_positions * axes + centre, unit.tet_indices
    mesh = ValidateTetMesh(verts, tets)
    rng = np.random.default_rng(seed)
    pos = jittered_cells(inside, lo, hi, dx, rng)
    material = _oven_material(poisson=0.45)
    samples = make_samples(mesh, pos, dx, material.density)
    porosity = rng.uniform(*porosity_range, size=len(pos))
    T0 = 298.15
    sim = SimulationConfig(grid_resolution=res, grid_spacing=dx, grid_origin=(0.0, 0.0, 0.0), time_step=dt,
                           gravity=(0.0, -9.8, 0.0), boundary_cells=3, device_name="auto")
    return SceneSpec(name, mesh, samples, sim, material, _oven_ding(0.05, p0, yield_stress=SCALED_YIELD_STRESS),
                     _oven_transport(T0, 473.15), steps=1000,
                     porosities=porosity.astype(np.float32), gas_moles=p0 * porosity / (gas_constant * T0),
                     notes=dict(shape=shape, axes=axes.tolist(), res=res, seed=seed, p0=p0))


def batched_doughs(count=48, seed=2, half_extent=(0.05, 0.15), max_particles=60000, dt=1.25e-5, p0=1000.0):
    """The batched workload: ``count`` independent doughs of uneven size (1k-60k particles each).

    Shapes and grids from ``dough_shapes(count, seed)``; dough i is seeded
    with ``default_rng([seed, i])``.  With the defaults: 48 doughs, ~1.0M
    particles, 64^3 and 128^3 grids, dt = 1.25e-5, 1000 substeps.
    """
    specs = []
    for i, d in enumerate(dough_shapes(count, seed, half_extent, max_particles)):
        specs.append(dough_scene(d["shape"], d["axes"], d["res"], seed=[seed, i], name=f"dough{i:02d}_{d['shape']}", dt=dt, p0=p0))
    return specs


def scaled(spec: SceneSpec, **sim_changes):
    return replace(spec, simulation=replace(spec.simulation, **sim_changes))


def build_simulator(spec: SceneSpec):
    """``MpmSimulator`` for a scene, with its per-particle initial fields applied."""
    import torch
    from .physics.mpm import MpmSimulator
    sim = MpmSimulator(spec.tet_mesh, spec.samples, spec.simulation, spec.material, spec.initial_velocity,
                       ding_config=spec.ding, transport_config=spec.transport)
    state = sim.current_state
    dev = sim.device_name
    if spec.water_masses is not None:
        state.transport_state.water_masses.copy_(torch.as_tensor(spec.water_masses, dtype=torch.float64, device=dev))
    if spec.porosities is not None:
        state.material_state.porosities.copy_(torch.as_tensor(spec.porosities, dtype=torch.float32, device=dev))
    if spec.gas_moles is not None:
        state.material_state.gas_moles.copy_(torch.as_tensor(spec.gas_moles, dtype=torch.float64, device=dev))
    if spec.velocities is not None:
        state.particle_velocities.copy_(torch.as_tensor(spec.velocities, dtype=torch.float32, device=dev))
    if spec.affine_velocities is not None:
        state.affine_velocities.copy_(torch.as_tensor(spec.affine_velocities, dtype=torch.float32, device=dev))
    if spec.temperatures is not None:
        state.transport_state.temperatures.copy_(torch.as_tensor(spec.temperatures, dtype=torch.float64, device=dev))
    return sim


def strain_rate_field(spec: SceneSpec, rates, centre=(0.5, 0.6, 0.5)):
    """Give a scene the affine velocity v = L (x - centre), C = L (reference tests/test_ding.py:124-127)."""
    L = np.asarray(rates, dtype=np.float32)
    x = spec.samples.particle_positions.astype(np.float32)
    return replace(spec, velocities=(x - np.asarray(centre, np.float32)) @ L.T,
                   affine_velocities=np.tile(L, (len(x), 1, 1)))


def demo_stretch(subdivisions=2, yield_stress=100.0, viscosity=1e8):
    """Ding dough cube in a diag(1.5, -0.75, 0.2) strain-rate flow (reference tests/test_ding.py:117-137)."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(subdivisions)
    samples = SampleParticles(mesh, 1000.0, 4)
    spec = SceneSpec("demo_stretch", mesh, samples, SimulationConfig(device_name="auto"), MaterialConfig(model="ding2019"),
                     DingConfig(yield_stress=yield_stress, viscosity=viscosity), TransportConfig(), steps=600)
    return strain_rate_field(spec, np.diag([1.5, -0.75, 0.2]))


def demo_gas(rate=0.005, dt=0.0005, subdivisions=1):
    """Gas-producing porous dough at 333.15 K (reference tests/test_gas.py:325-333)."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(subdivisions)
    ding = DingConfig(gas_enabled=True, porosity=0.2, yield_stress=20., viscosity=1e5, gas_initial_pressure_pa=1000.,
                      gas_ambient_pressure_pa=1000., gas_production_rate=rate)
    transport = TransportConfig(enabled=True, water_enabled=False, initial_temperature_k=333.15, conductivity=0.,
                                heat_transfer=0., tray_heat_transfer=0.)
    sim = SimulationConfig(device_name="auto", time_step=dt, gravity=(0.0, 0.0, 0.0))
    return SceneSpec("demo_gas", mesh, SampleParticles(mesh, 1000., 4), sim,
                     MaterialConfig(model="ding2019", young_modulus=10000.), ding, transport, steps=100)


def demo_heated(viscosity=1e8, yield_stress=100.0, dt=1e-4):
    """Temperature-dependent porous dough with transport in a strain-rate flow (reference tests/test_ding_thermal.py:212-225)."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(1)
    sim = SimulationConfig(device_name="auto", time_step=dt, gravity=(0.0, 0.0, 0.0))
    spec = SceneSpec("demo_heated", mesh, SampleParticles(mesh, 1000., 4), sim, MaterialConfig(model="ding2019"),
                     DingConfig(temperature_enabled=True, porosity=.2, yield_stress=yield_stress, viscosity=viscosity),
                     TransportConfig(enabled=True, water_enabled=False, initial_temperature_k=300.), steps=100)
    return strain_rate_field(spec, np.diag([1.5, -.75, .2]))


def demo_autodiff(velocity=0.2):
    """Elastic demo cube on a 20^3 grid (dx 0.05, dt 1e-3, no gravity) moving at ``velocity`` along x
    (reference tests/test_warp_autodiff.py:24-33)."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(1)
    samples = SampleParticles(mesh, 1000.0, 4)
    sim = SimulationConfig(grid_resolution=20, grid_spacing=0.05, time_step=0.001, gravity=(0.0, 0.0, 0.0), device_name="auto")
    n = len(samples.particle_positions)
    v = np.zeros((n, 3), np.float32)
    v[:, 0] = velocity
    return SceneSpec("demo_autodiff", mesh, samples, sim, MaterialConfig(), DingConfig(), TransportConfig(), steps=3, velocities=v)


def demo_viscous_gradient(viscosity=1000.0, yield_stress=5.0, porosity=0.0):
    """Ding demo cube (dt 1e-3, no gravity) in a diag(1.5, -0.75, 0.2) strain-rate flow
    (reference tests/test_ding.py:142-151)."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(1)
    sim = SimulationConfig(time_step=0.001, gravity=(0.0, 0.0, 0.0), device_name="auto")
    spec = SceneSpec("demo_viscous_gradient", mesh, SampleParticles(mesh, 1000.0, 4), sim, MaterialConfig(model="ding2019"),
                     DingConfig(yield_stress=yield_stress, viscosity=viscosity, porosity=porosity), TransportConfig(), steps=5)
    return strain_rate_field(spec, np.diag([1.5, -0.75, 0.2]))


def demo_transport(heat_transfer=20.0, gas=True, water_stress=0.0, spacing=0.025, dt=1e-4, subdivisions=2):
    """Ding dough cube with full heat and water transport (reference tests/test_transport.py:12-20 with the
    TransportConfig of its heat-transfer gradient test, 96-121): 40^3 grid, dt 1e-4, no gravity, pore gas at
    porosity 0.2, temperature-dependent yield stress and viscosity."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(subdivisions)
    sim = SimulationConfig(time_step=dt, gravity=(0.0, 0.0, 0.0), grid_spacing=spacing, grid_resolution=int(round(1 / spacing)),
                           device_name="auto")
    ding = DingConfig(gas_enabled=gas, porosity=0.2 if gas else 0.0, gas_production_rate=0.001, temperature_enabled=True)
    transport = TransportConfig(enabled=True, heat_transfer=heat_transfer, water_stress=water_stress)
    return SceneSpec("demo_transport", mesh, SampleParticles(mesh, 1000.0, 4), sim, MaterialConfig(model="ding2019", young_modulus=10000.0),
                     ding, transport, steps=4)


def demo_resting(heat_transfer=20.0, settling_duration_s=0.0004):
    """Fixed-duration rest before baking (reference tests/test_settling.py:8-12): the transport cube of
    tests/test_transport.py:12-20 with pore gas, default water stress, gravity -9.81 and a 0.4 ms rest
    (4 substeps at dt 1e-4)."""
    spec = demo_transport(heat_transfer=heat_transfer, gas=True, water_stress=TransportConfig().water_stress)
    spec.name = "demo_resting"
    spec.simulation = replace(spec.simulation, gravity=(0.0, -9.81, 0.0))
    spec.transport = replace(spec.transport, settling_duration_s=settling_duration_s)
    spec.steps = 8
    return spec


def demo_auto_settling(heat_transfer=20.0, gravity=(0.0, 0.0, 0.0)):
    """Automatic quasi-static detection (reference tests/test_auto_settling.py:9-16): at most 2 ms of rest,
    0.2 ms minimum, 0.3 ms hold, speed tolerance 1e-4 m/s and rate tolerance 0.01 1/s."""
    spec = demo_transport(heat_transfer=heat_transfer, gas=True, water_stress=TransportConfig().water_stress)
    spec.name = "demo_auto_settling"
    spec.simulation = replace(spec.simulation, gravity=tuple(gravity))
    spec.transport = replace(spec.transport, settling_duration_s=0.002, settling_auto=True, settling_min_duration_s=0.0002,
                             settling_hold_duration_s=0.0003, settling_speed_tolerance=1e-4, settling_rate_tolerance=0.01)
    spec.steps = 8
    return spec


def demo_translation(velocity=(0.2, -0.1, 0.05), gravity=(0.0, 0.0, 0.0), time_step=1e-4):
    """Neo-Hookean demo cube translating freely (reference tests/test_mpm.py:10-15): CreateDemoMesh(3), 4
    particles per tet at density 1000, the default 64^3 grid."""
    from .geometry.mesh_loader import CreateDemoMesh
    from .geometry.particle_sampling import SampleParticles
    mesh = CreateDemoMesh(3)
    sim = SimulationConfig(gravity=tuple(gravity), time_step=time_step, device_name="auto")
    return SceneSpec("demo_translation", mesh, SampleParticles(mesh, 1000.0, 4), sim, MaterialConfig(), DingConfig(),
                     TransportConfig(), steps=30, initial_velocity=tuple(velocity))
