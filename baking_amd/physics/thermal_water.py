"""Heat and water transport of the baking dough (reference ``physics/thermal_water.py:1-367``).

One explicit split per substep: mesh reconstruction and boundary-face areas ->
scalar P2G of volume, capacity, energy, water, Robin heat, water loss and
exchange -> six-neighbour finite-volume heat and water fluxes with Robin terms
-> scalar G2P -> local gas source.  Water stress -c_hat (s - s0) J I is added
to the mechanical stress; saturations are refreshed after the mechanical step.

The fused substep (``MpmSimulator.Step``) runs all of this inside
``bk_substep``; the methods of :class:`ThermalWaterTransport` and the kernels
below are the unfused entries with the reference's signatures.
"""
from dataclasses import dataclass, replace

import numpy as np
import torch

from ..launch import Kernel
from .materials.gas import gas_constant  # noqa: F401  (re-exported like the reference)

SurfaceAreas = Kernel("bk_SurfaceAreas", ["f32", "i32", "i32", "f64", "f64", "f", "f"])
ScalarToGrid = Kernel("bk_ScalarToGrid", ["f32", "f32", "f32", "f64", "f64", "f64", "f64", "f64", "f32", "f", "f", "f", "f", "f", "f",
                                          "i"] + ["f64"] * 9 + ["v3", "f", "i"])
GridFields = Kernel("bk_GridFields", ["f64"] * 6 + ["f"], dim64=True)
DiffuseGrid = Kernel("bk_DiffuseGrid", ["f64"] * 10 + ["i32", "f", "f", "f", "i", "f", "i", "f"], dim64=True)
ScalarToParticles = Kernel("bk_ScalarToParticles", ["f32", "f64", "f64", "f64", "f64", "f64", "f32", "f32", "f64", "f64", "f64", "f64",
                                                    "f64", "f64", "f32", "f64", "f64", "f64", "f32", "f32", "f64", "f", "v3", "f", "i",
                                                    "f"])
LocalGas = Kernel("bk_LocalGas", ["f64", "f64", "f64", "f64", "f64", "f32", "f32", "f32", "f", "f", "f", "f", "i"])
WaterStress = Kernel("bk_WaterStress", ["f32", "f64", "f32", "f32", "f", "f"])
Saturations = Kernel("bk_Saturations", ["f64", "f32", "f32", "f64", "f"])


@dataclass
class TransportState:
    temperatures: object
    water_masses: object
    saturations: object
    water_losses: object
    heat_exchanges: object
    particle_masses: object


def BoundaryFaces(tet_mesh, samples):
    """Outward boundary triangles and, per face, the owning particle (thermal_water.py:51-74).

    A face seen once is on the boundary; it is wound outward relative to its
    tet's opposite vertex, and owned by the particle of that tet closest to the
    face centre (the first such particle on a tie).  Face order is first-seen
    order over tets and local faces.  Vectorised over all 4T faces (a sort of
    the face keys finds the faces seen once), with the reference's arithmetic
    per face, so the output equals its loop's.
    """
    tets = np.asarray(tet_mesh.tet_indices)
    verts = np.asarray(tet_mesh.vertex_positions, dtype=np.float64)
    ntet = len(tets)
    keep = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])  # np.delete(tet, opposite) for opposite 0..3
    faces = tets[:, keep].reshape(-1, 3).copy()                      # face 4 t + o
    opposite = tets.reshape(-1)                                      # the vertex each face leaves out
    keys = np.sort(faces, axis=1)
    _, inverse, seen = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    boundary = np.flatnonzero(seen[inverse.reshape(-1)] == 1)
    faces, opposite = faces[boundary], opposite[boundary]
    owner_tet = boundary // 4
    pts = verts[faces]
    normal = np.cross(pts[:, 1] - pts[:, 0], pts[:, 2] - pts[:, 0])
    outward = np.einsum("ij,ij->i", normal, verts[opposite] - pts[:, 0]) > 0
    faces[outward] = faces[outward][:, [0, 2, 1]]
    # owners: nearest particle of the face's tet to the face centre (first on ties)
    source = np.asarray(samples.source_tets)
    order = np.argsort(source, kind="stable")
    first = np.searchsorted(source[order], np.arange(ntet))
    count = np.bincount(source, minlength=ntet)
    members_per_face = count[owner_tet]
    face_of = np.repeat(np.arange(len(faces)), members_per_face)
    rank = np.arange(len(face_of)) - np.repeat(np.cumsum(members_per_face) - members_per_face, members_per_face)
    member = order[first[owner_tet][face_of] + rank]
    centre = verts[faces].mean(axis=1)
    dist = np.linalg.norm(np.asarray(samples.particle_positions)[member] - centre[face_of], axis=1)
    best = np.lexsort((rank, dist, face_of))
    head = np.concatenate(([True], face_of[best][1:] != face_of[best][:-1]))
    owners = member[best][head]
    return faces.astype(np.int32).reshape(-1, 3), owners.astype(np.int32)


def SurfaceSlots(owners, particle_count):
    """Group faces by owner particle: slots (one per distinct owner) with their faces in face order.

    Summing a slot's face areas in face order reproduces the reference's serial
    per-owner atomic sums.
    """
    owners = np.asarray(owners, dtype=np.int64)
    slot_owner = np.unique(owners)
    slot_of = np.full(particle_count, -1, dtype=np.int32)
    slot_of[slot_owner] = np.arange(len(slot_owner), dtype=np.int32)
    face_slot = slot_of[owners]
    order = np.argsort(face_slot, kind="stable")
    counts = np.bincount(face_slot, minlength=len(slot_owner))
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    return dict(slot_owner=slot_owner.astype(np.int32), particle_slot=slot_of, slot_face_offsets=offsets,
                slot_faces=order.astype(np.int32))


class ThermalWaterTransport:
    def __init__(self, simulator, config):
        self.simulator = simulator
        self.config = config
        self.device_name = simulator.device_name
        self.requires_grad = simulator.retain_history
        samples = simulator.particle_samples
        density = simulator.material.material_config.density
        dev = self.device_name
        volumes = np.asarray(samples.reference_volumes, dtype=np.float64)
        self.reference_volumes = torch.as_tensor(volumes.astype(np.float32), device=dev)
        self.dry_masses = torch.as_tensor((density - config.water_density * config.initial_saturation) * volumes, device=dev)
        self.capacities = torch.as_tensor(density * config.specific_heat * volumes, device=dev)
        self.heat_transfer_parameter = torch.tensor([config.heat_transfer], dtype=torch.float32, device=dev)
        faces, owners = BoundaryFaces(simulator.tet_mesh, samples)
        self.faces = torch.as_tensor(faces, device=dev)
        self.owners = torch.as_tensor(owners, device=dev)
        slots = SurfaceSlots(owners, len(volumes))
        self.slots = {k: torch.as_tensor(v, device=dev) for k, v in slots.items()}
        self.error_flags = torch.zeros(2, dtype=torch.int32, device=dev)
        self.workspace = None

    def Array(self, count, dtype=torch.float64):
        return torch.zeros(count, dtype=dtype, device=self.device_name)

    def InitializeState(self, state):
        count = len(state.particle_positions)
        c = self.config
        volumes = torch.as_tensor(np.asarray(self.simulator.particle_samples.reference_volumes, dtype=np.float64),
                                  device=self.device_name)
        state.transport_state = TransportState(
            torch.full((count,), c.initial_temperature_k, dtype=torch.float64, device=self.device_name),
            c.water_density * c.initial_saturation * volumes,
            torch.full((count,), c.initial_saturation, dtype=torch.float64, device=self.device_name),
            self.Array(count), self.Array(count),
            torch.as_tensor(np.asarray(self.simulator.particle_samples.particle_masses, dtype=np.float32),
                            device=self.device_name))
        self.error_flags.zero_()

    def Advance(self, state, time_step, settling=False):
        """Unfused transport pre-pass from ``state`` (thermal_water.py:296-355); returns the prepared state."""
        c = self.config
        sim = self.simulator.simulation_config
        count = len(state.particle_positions)
        nodes = sim.grid_resolution ** 3

        class Workspace:
            pass

        w = Workspace()
        w.heat_parameter = self.Array(1, torch.float32) if settling else self.heat_transfer_parameter
        w.vertices = torch.zeros(len(self.simulator.tet_mesh.vertex_positions), 3, dtype=torch.float32, device=self.device_name)
        self.simulator.mesh_binding.Reconstruct(state, w.vertices)
        w.exposed = self.Array(count)
        w.contact = self.Array(count)
        floor = sim.grid_origin[1] + (sim.boundary_cells - 1) * sim.grid_spacing
        SurfaceAreas(len(self.faces), w.vertices, self.faces, self.owners, w.exposed, w.contact, floor, c.contact_distance)
        for name in ("volumes", "capacities", "energy", "water", "heat", "loss", "exchange", "temperatures", "saturations",
                     "temperature_changes", "water_changes"):
            setattr(w, name, self.Array(nodes))
        w.particle_heat = self.Array(count)
        w.particle_loss = self.Array(count)
        old = state.transport_state
        out = TransportState(self.Array(count), self.Array(count), self.Array(count), self.Array(count), self.Array(count),
                             self.Array(count, torch.float32))
        material_temperatures = self.Array(count, torch.float32)
        densities = self.Array(count, torch.float32)
        origin = tuple(sim.grid_origin)
        ScalarToGrid(count, state.particle_positions, state.deformation_gradients, self.reference_volumes, self.capacities,
                     old.temperatures, old.water_masses, w.exposed, w.contact, w.heat_parameter, c.oven_temperature_k,
                     c.tray_temperature_k, 0.0 if settling else c.tray_heat_transfer, 0.0 if settling else c.surface_water_loss,
                     c.water_density, c.latent_heat, int(c.water_enabled), w.volumes, w.capacities, w.energy, w.water, w.heat,
                     w.loss, w.exchange, w.particle_heat, w.particle_loss, origin, sim.grid_spacing, sim.grid_resolution)
        GridFields(nodes, w.volumes, w.capacities, w.energy, w.water, w.temperatures, w.saturations, c.water_density)
        DiffuseGrid(nodes, w.volumes, w.capacities, w.temperatures, w.saturations, w.water, w.heat, w.loss, w.exchange,
                    w.temperature_changes, w.water_changes, self.error_flags, 0.0 if settling else c.conductivity,
                    0.0 if settling else c.water_diffusivity, c.water_density, int(c.water_enabled), sim.grid_spacing,
                    sim.grid_resolution, time_step)
        flags = self.error_flags.cpu().tolist()
        if flags[0] or flags[1]:
            raise RuntimeError('Heat/water explicit stability bound exceeded; reduce time_step; '
                               'no temperature or water clipping is applied')
        ScalarToParticles(count, state.particle_positions, old.temperatures, old.water_masses, old.water_losses,
                          old.heat_exchanges, self.dry_masses, self.reference_volumes, state.deformation_gradients, w.water,
                          w.temperature_changes, w.water_changes, w.particle_heat, w.particle_loss, out.temperatures,
                          material_temperatures, out.water_masses, out.water_losses, out.heat_exchanges, out.particle_masses,
                          densities, out.saturations, c.water_density, origin, sim.grid_spacing, sim.grid_resolution, time_step)
        material_state = replace(state.material_state, temperatures=material_temperatures, reference_densities=densities)
        material = self.simulator.material
        if material.ding_config.gas_enabled:
            material_state.gas_moles = self.Array(count)
            material_state.gas_pressures = self.Array(count)
            LocalGas(count, state.material_state.gas_moles, old.temperatures, out.temperatures, material_state.gas_moles,
                     material_state.gas_pressures, state.deformation_gradients, state.material_state.porosities,
                     material.production_parameter, material.material_config.density,
                     material.ding_config.gas_reaction_temperature_k, material.ding_config.gas_reaction_width_k, time_step,
                     int(not settling))
        w.prepared = replace(state, material_state=material_state, transport_state=out)
        self.workspace = w
        return w.prepared

    def AddWaterStress(self, state, stress):
        out = torch.zeros_like(stress)
        WaterStress(len(out), state.deformation_gradients, state.transport_state.saturations, stress, out,
                    self.config.initial_saturation, self.config.water_stress if self.config.water_enabled else 0.0)
        if self.workspace is not None:
            self.workspace.stress_before_water = stress
        return out

    def FinishStep(self, prepared, output):
        output.transport_state = replace(prepared.transport_state, saturations=self.Array(len(output.particle_positions)))
        Saturations(len(output.particle_positions), output.transport_state.water_masses, self.reference_volumes,
                    output.deformation_gradients, output.transport_state.saturations, self.config.water_density)
