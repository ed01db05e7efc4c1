"""Background grid (reference ``physics/grid.py:4-38``).

The fused substep stores the grid tiled: 4x4x4 node blocks contiguous, so one
wavefront reads one block with unit stride.  ``node_masses``,
``node_momenta`` and ``node_velocities`` give the reference's dense layout
(node id (i*res + j)*res + k) as fresh tensors.
"""
import torch

from .. import _native
from ..launch import Kernel

UpdateGrid = Kernel("bk_UpdateGrid", ["f32", "f32", "f32", "i", "i", "v3", "f"], dim64=True)


class GridState:
    def __init__(self, simulation_config, device_name, requires_grad=False):
        res = simulation_config.grid_resolution
        self.resolution = res
        self.node_count = res ** 3
        self.blocks_per_axis = (res + 3) // 4
        self.tiled_count = self.blocks_per_axis ** 3 * 64
        self.device_name = device_name
        n = self.tiled_count
        self.tiled_masses = torch.zeros(n, dtype=torch.float32, device=device_name)
        self.tiled_momenta = torch.zeros(n, 3, dtype=torch.float32, device=device_name)
        self.tiled_velocities = torch.zeros(n, 3, dtype=torch.float32, device=device_name)

    def Clear(self):
        self.tiled_masses.zero_()
        self.tiled_momenta.zero_()
        self.tiled_velocities.zero_()

    def _untile(self, tiled, comps):
        out = torch.empty((self.node_count, comps) if comps > 1 else (self.node_count,), dtype=torch.float32,
                          device=self.device_name)
        _native.check(_native.load().bk_grid_untile(_native.ptr(tiled), _native.ptr(out), self.resolution, comps,
                                                    _native.stream_ptr()), "bk_grid_untile")
        return out

    @property
    def node_masses(self):
        return self._untile(self.tiled_masses, 1)

    @property
    def node_momenta(self):
        return self._untile(self.tiled_momenta, 3)

    @property
    def node_velocities(self):
        return self._untile(self.tiled_velocities, 3)
