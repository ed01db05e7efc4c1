"""``launch(kernel, dim, inputs)``: the drop-in for the reference's ``wp.launch``.

Each reference Warp kernel has a :class:`Kernel` of the same name in the same
module path (``baking_amd.physics.transfer.ParticleToGrid`` ...).  Inputs keep
the Warp kernel's order; arrays are contiguous torch tensors on the ROCm
device (vec3 -> (N,3) float32, mat33 -> (N,3,3) float32, wp.float64 ->
float64, int -> int32); wp.vec3/wp.vec4 values are Python sequences.
"""
from __future__ import annotations

from ctypes import c_float, c_int32, c_int64

import torch

from . import _native

_DTYPES = {"f32": torch.float32, "f64": torch.float64, "i32": torch.int32}


class Kernel:
    """One C-ABI entry ``symbol`` with argument kinds matching the Warp kernel's inputs.

    kinds: 'f32'/'f64'/'i32' device arrays, 'v3'/'v4' host vectors, 'f' float, 'i' int.
    """

    def __init__(self, symbol, kinds, dim64=False):
        self.symbol = symbol
        self.kinds = list(kinds)
        self.dim64 = dim64
        self.__name__ = symbol[3:]

    def __call__(self, dim, *inputs, stream=None):
        if len(inputs) != len(self.kinds):
            raise TypeError(f"{self.__name__} takes {len(self.kinds)} inputs, got {len(inputs)}")
        _native.require_gpu()
        lib = _native.load()
        args, keep = [], []
        for index, (kind, value) in enumerate(zip(self.kinds, inputs)):
            if kind in _DTYPES:
                if not isinstance(value, torch.Tensor) or value.dtype != _DTYPES[kind] or not value.is_cuda:
                    raise TypeError(f"{self.__name__} input {index} must be a {kind} tensor on the GPU")
                args.append(_native.ptr(value))
            elif kind in ("v3", "v4"):
                values = [float(v) for v in value]
                if len(values) != int(kind[1]):
                    raise TypeError(f"{self.__name__} input {index} must have {kind[1]} components")
                p, arr = _native.host_f32(values)
                keep.append(arr)
                args.append(p)
            elif kind == "f":
                args.append(c_float(float(value)))
            elif kind == "i":
                args.append(c_int32(int(value)))
            else:
                raise ValueError(kind)
        args.append(c_int64(int(dim)) if self.dim64 else c_int32(int(dim)))
        args.append(_native.stream_ptr(stream))
        _native.check(getattr(lib, self.symbol)(*args), self.__name__)


def launch(kernel, dim, inputs=(), device=None, stream=None):
    """Enqueue ``kernel`` over ``dim`` items on the current stream (``wp.launch`` signature)."""
    kernel(dim, *inputs, stream=stream)
