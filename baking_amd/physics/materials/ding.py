"""Ding-2019 viscoplastic porous dough (reference ``physics/materials/ding.py:1-327``).

The return map (fp64 SVD of the trial elastic gradient, Hencky strains, 32
Newton steps on the scalar overstress, coupled end-state porosity when gas is
on) runs on the GPU in ``UpdateDing``; the fused substep calls the same device
code from its G2P pass.  Stored state stays float32 as in the reference.
"""
import math
from dataclasses import dataclass

import numpy as np
import torch

from ... import _native
from ...launch import Kernel
from .elastic import ElasticMaterial
from .gas import CombineGasStress, RefreshGas, gas_constant

UpdateDing = Kernel("bk_UpdateDing",
                    ["f32", "f32", "i32", "f32", "f32", "f32", "f32", "f32", "f32", "f32", "f32", "i32", "f32", "f32", "f32",
                     "i32", "f32", "i", "f", "f", "f", "f", "i", "v4", "v4", "v4", "f", "f32"])
_PorosityCoefficient = Kernel("bk_porosity_coefficient", ["f32", "f32"])


@dataclass
class DingState:
    """Gas content is mol / initial m^3; pressures are absolute Pa."""
    elastic_gradients: object
    plastic_gradients: object
    kirchhoff_stresses: object
    return_residuals: object
    porosities: object
    temperatures: object
    gelled: object
    yield_stresses: object
    viscosities: object
    gas_moles: object
    gas_pressures: object
    porosity_residuals: object
    reference_densities: object


def TemperatureFactor(temperature_value, temperature_knots, factor_values):
    """Piecewise-linear factor with constant end extrapolation (ding.py:51-53)."""
    return float(np.interp(temperature_value, temperature_knots, factor_values))


def porosity_coefficient(porosities):
    """a(n) = 9 / (4 ln^2(1/n) (3.5 n + 1)), a(0) = 0, evaluated on the GPU (ding.py:68-74)."""
    out = torch.empty_like(porosities)
    _PorosityCoefficient(porosities.numel(), porosities, out)
    return out


class DingMaterial(ElasticMaterial):
    def __init__(self, material_config, ding_config, device_name, requires_grad=False, initial_temperature_k=293.15):
        super().__init__(material_config)
        ding_config.Validate()
        self.initial_temperature_k = initial_temperature_k
        self.ding_config = ding_config
        self.device_name = device_name
        self.viscosity_parameter = torch.tensor([ding_config.viscosity], dtype=torch.float32, device=device_name)
        self.error_flags = torch.zeros(3, dtype=torch.int32, device=device_name)
        self.production_parameter = torch.tensor([ding_config.gas_production_rate], dtype=torch.float32, device=device_name)
        n0 = ding_config.porosity
        self.pressure_coefficient = 0.0 if n0 == 0 else 9.0 / (4.0 * math.log(1.0 / n0) ** 2 * (3.5 * n0 + 1.0))
        self.Reset()

    def Reset(self):
        self.error_flags.zero_()

    def InitializeState(self, particle_state):
        """Fe = Fp = I, n = n0, N0 = p0 n0 / (R T0), tables at T0 (ding.py:243-269)."""
        count = len(particle_state.particle_positions)
        device = particle_state.particle_positions.device
        c = self.ding_config
        T0 = self.initial_temperature_k
        gelled = int(c.temperature_enabled and T0 >= c.temperature_knots_k[3])
        yf = TemperatureFactor(T0, c.temperature_knots_k, c.yield_factors) if c.temperature_enabled else 1.0
        vf = TemperatureFactor(T0, c.temperature_knots_k, c.viscosity_factors) if c.temperature_enabled else 1.0
        eye = torch.eye(3, dtype=torch.float32, device=device).expand(count, 3, 3).contiguous()

        def full(value, dtype=torch.float32):
            return torch.full((count,), value, dtype=dtype, device=device)

        visc = float(self.viscosity_parameter.cpu()[0])
        moles = c.gas_initial_pressure_pa * c.porosity / (gas_constant * T0) if c.gas_enabled else 0.0
        particle_state.material_state = DingState(
            elastic_gradients=eye.clone(), plastic_gradients=eye.clone(),
            kirchhoff_stresses=torch.zeros(count, 3, 3, dtype=torch.float32, device=device),
            return_residuals=full(0.0), porosities=full(c.porosity), temperatures=full(T0),
            gelled=full(gelled, torch.int32), yield_stresses=full(c.yield_stress * yf),
            viscosities=full(float(np.float32(visc)) * vf), gas_moles=full(moles, torch.float64),
            gas_pressures=full(c.gas_initial_pressure_pa if c.gas_enabled else 0.0, torch.float64),
            porosity_residuals=full(0.0), reference_densities=full(self.material_config.density))

    def ComputeStress(self, particle_state, kirchhoff_stresses, error_flags):
        """Total Kirchhoff stress: solid + gas pore pressure (ding.py:271-283)."""
        state = particle_state.material_state
        if not self.ding_config.gas_enabled:
            kirchhoff_stresses.copy_(state.kirchhoff_stresses)
            return
        sim = getattr(self, "simulation_config", None)
        CombineGasStress(len(kirchhoff_stresses), state.kirchhoff_stresses, particle_state.deformation_gradients,
                         state.porosities, state.gas_moles, state.temperatures, kirchhoff_stresses, self.error_flags,
                         self.ding_config.gas_ambient_pressure_pa, self.ding_config.porosity, state.reference_densities,
                         self.material_config.lame_lambda + 2 * self.material_config.shear_modulus,
                         sim.time_step if sim else 0.0, sim.grid_spacing if sim else 1.0, sim.cfl_limit if sim else 0.4)
        self.CheckState()

    def GasWaveSpeed(self, particle_state):
        internal = particle_state.material_state
        porosities = internal.porosities.cpu().numpy().astype(np.float64)
        pressures = internal.gas_pressures.cpu().numpy()
        J = np.linalg.det(particle_state.deformation_gradients.cpu().numpy().astype(np.float64))
        bulk = pressures * (1 + (1 - self.ding_config.porosity + porosities) / porosities)
        modulus = self.material_config.lame_lambda + 2 * self.material_config.shear_modulus
        return float(np.sqrt((modulus + bulk) * J / internal.reference_densities.cpu().numpy()).max())

    def RaiseFor(self, flags):
        """Raise the reference's error for Ding flags [domain, coupled, gas CFL] (ding.py:294-301)."""
        if flags[2]:
            raise RuntimeError("Gas acoustic CFL exceeds the explicit limit; reduce time_step. Pressure is not capped")
        if flags[1]:
            raise RuntimeError("Coupled porosity return has no resolved interior solution; reduce dt/loading. "
                               "No pore-volume clipping or closed-pore law is applied")
        if flags[0]:
            raise RuntimeError("Ding porosity left [0, 1); no constitutive clamping is applied. Reduce dt or loading; "
                               "the paper gives no closure beyond this domain")

    def CheckState(self):
        self.RaiseFor(self.error_flags.cpu().tolist())

    def AdvanceState(self, input_state, output_state, time_step):
        """UpdateDing, density copy and gas refresh from input to output (ding.py:303-327)."""
        if output_state.material_state is None:
            self.InitializeState(output_state)
        old = input_state.material_state
        new = output_state.material_state
        c = self.ding_config
        n = len(output_state.particle_positions)
        UpdateDing(n, old.elastic_gradients, old.porosities, old.gelled, output_state.deformation_gradients,
                   output_state.affine_velocities, new.elastic_gradients, new.plastic_gradients, new.kirchhoff_stresses,
                   new.return_residuals, new.porosities, new.temperatures, new.gelled, new.yield_stresses, new.viscosities,
                   self.viscosity_parameter, self.error_flags, new.porosity_residuals, int(c.gas_enabled),
                   self.material_config.shear_modulus, self.material_config.lame_lambda, c.yield_stress, c.porosity,
                   int(c.temperature_enabled), c.temperature_knots_k, c.yield_factors, c.viscosity_factors, time_step,
                   old.temperatures)
        new.reference_densities.copy_(old.reference_densities)
        if c.gas_enabled:
            RefreshGas(n, old.gas_moles, new.gas_moles, output_state.deformation_gradients, new.porosities,
                       new.gas_pressures, new.temperatures)
