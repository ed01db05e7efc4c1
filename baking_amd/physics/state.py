"""Particle state containers (reference ``physics/state.py:6-46``).

Arrays are torch tensors on the simulator's ROCm device: vec3 -> (N, 3)
float32, mat33 -> (N, 3, 3) float32, row-major like Warp's mat33.
"""
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class ParticleState:
    particle_positions: object
    particle_velocities: object
    deformation_gradients: object
    affine_velocities: object
    material_state: object = None
    transport_state: object = None


@dataclass
class ParticleData:
    particle_masses: object
    reference_volumes: object
    particle_count: int


def AllocateState(particle_count, device_name, requires_grad=False):
    """Zero-filled x, v, F, C for ``particle_count`` particles."""
    def zeros(*shape):
        return torch.zeros(*shape, dtype=torch.float32, device=device_name, requires_grad=requires_grad)
    return ParticleState(zeros(particle_count, 3), zeros(particle_count, 3), zeros(particle_count, 3, 3),
                         zeros(particle_count, 3, 3))


def CreateParticleState(particle_samples, device_name, initial_velocity=(0.0, 0.0, 0.0), requires_grad=False):
    """State at rest shape (F = I, C = 0) with a uniform initial velocity."""
    count = len(particle_samples.particle_positions)
    velocity = np.asarray(initial_velocity, dtype=np.float32)
    if velocity.shape != (3,) or not np.isfinite(velocity).all():
        raise ValueError("initial_velocity must have three finite components")
    state = AllocateState(count, device_name, requires_grad)
    with torch.no_grad():
        state.particle_positions.copy_(torch.from_numpy(particle_samples.particle_positions.astype(np.float32)))
        state.particle_velocities.copy_(torch.from_numpy(np.tile(velocity, (count, 1))))
        state.deformation_gradients.copy_(torch.eye(3).expand(count, 3, 3))
    data = ParticleData(torch.as_tensor(np.asarray(particle_samples.particle_masses, dtype=np.float32), device=device_name),
                        torch.as_tensor(np.asarray(particle_samples.reference_volumes, dtype=np.float32), device=device_name),
                        count)
    return state, data
