"""Ideal-gas pore pressure (reference ``physics/materials/gas.py:1-45``).

Gas content is mol per initial m^3; pressures are absolute Pa.
"""
import torch

from ... import _native
from ...launch import Kernel

gas_constant = 8.31446261815324

CombineGasStress = Kernel("bk_CombineGasStress",
                          ["f32", "f32", "f32", "f64", "f32", "f32", "i32", "f", "f", "f32", "f", "f", "f", "f"])
RefreshGas = Kernel("bk_RefreshGas", ["f64", "f64", "f32", "f32", "f64", "f32"])
_GasPressure = Kernel("bk_gas_pressure", ["f64", "f32", "f32", "f32", "f64"])


def gas_pressure(moles, temperature, volume_ratio, porosity):
    """Element-wise p = N R T / (J n) on the GPU (gas.py:11-13): float64 moles, float32 T, J, n."""
    out = torch.empty_like(moles, dtype=torch.float64)
    _GasPressure(moles.numel(), moles, temperature, volume_ratio, porosity, out)
    return out
