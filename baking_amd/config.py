"""Scene parameters in SI units (metre, kilogram, second, pascal, kelvin).

The four parameter groups the substep reads -- material, simulation (grid and time step), Ding dough and
heat/water transport -- as frozen dataclasses with the reference's field names and defaults
(``config.py:10-75, 109-196``).  ``Validate`` walks a table of rules per group, each a condition the
parameters must satisfy and the message raised (the reference's texts, in its order, so the first
violated rule is the one the reference reports).  A non-finite value fails the rule that reads it.
The reference's INI loader and its experiment/CLI sections are not part of this package.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass
from typing import Callable, NamedTuple


class Rule(NamedTuple):
    """``holds(config, context)`` must be true; otherwise ``message`` (text, or a function of the config) is raised."""

    holds: Callable
    message: object


def check_rules(config, rules, context=None):
    for rule in rules:
        if not rule.holds(config, context):
            text = rule.message(config) if callable(rule.message) else rule.message
            raise ValueError(text)


def _values(config, names):
    return [getattr(config, name) for name in names]


def positive(*names):
    """Every named field is finite and > 0."""
    return lambda c, _: all(math.isfinite(v) and v > 0.0 for v in _values(c, names))


def nonnegative(*names):
    """Every named field is finite and >= 0."""
    return lambda c, _: all(math.isfinite(v) and v >= 0.0 for v in _values(c, names))


def finite(*names):
    return lambda c, _: all(math.isfinite(v) for v in _values(c, names))


def finite_vectors(*names):
    return lambda c, _: all(math.isfinite(v) for name in names for v in getattr(c, name))


@dataclass(frozen=True)
class MaterialConfig:
    """Bulk elastic material (reference ``config.py:10-38``)."""

    density: float = 1000.0
    young_modulus: float = 20000.0
    poisson_ratio: float = 0.3
    model: str = "neo_hookean"

    def Validate(self):
        check_rules(self, MATERIAL_RULES)

    @property
    def shear_modulus(self):
        return self.young_modulus / (2.0 * (1.0 + self.poisson_ratio))

    @property
    def lame_lambda(self):
        nu = self.poisson_ratio
        return self.young_modulus * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))

    @property
    def wave_speed(self):
        """Elastic P-wave speed sqrt((lambda + 2 mu) / rho)."""
        return math.sqrt((self.lame_lambda + 2.0 * self.shear_modulus) / self.density)


MATERIAL_RULES = (
    Rule(positive("density"), "density must be positive and finite"),
    Rule(positive("young_modulus"), "young_modulus must be positive and finite"),
    Rule(lambda c, _: c.model in ("neo_hookean", "ding2019"), "material model must be neo_hookean or ding2019"),
    Rule(lambda c, _: -1.0 < c.poisson_ratio < 0.5, "poisson_ratio must be in (-1, 0.5)"),
)


@dataclass(frozen=True)
class SimulationConfig:
    """Background grid and time stepping (reference ``config.py:41-75``)."""

    grid_resolution: int = 40
    grid_spacing: float = 0.025
    grid_origin: tuple = (0.0, 0.0, 0.0)
    time_step: float = 1.0e-4
    gravity: tuple = (0.0, -9.81, 0.0)
    boundary_cells: int = 3
    steps_per_frame: int = 40
    device_name: str = "auto"
    cfl_limit: float = 0.4

    def Validate(self, material_config):
        material_config.Validate()
        check_rules(self, SIMULATION_RULES, material_config)


SIMULATION_RULES = (
    Rule(lambda c, _: c.grid_resolution >= 8 and c.boundary_cells >= 2,
         "grid_resolution >= 8 and boundary_cells >= 2 are required"),
    Rule(lambda c, _: 2 * c.boundary_cells + 2 < c.grid_resolution, "boundary layers leave no usable interior"),
    Rule(positive("grid_spacing"), "grid_spacing must be positive and finite"),
    Rule(positive("time_step"), "time_step must be positive and finite"),
    Rule(lambda c, _: c.steps_per_frame >= 1, "steps_per_frame must be positive"),
    Rule(lambda c, _: len(c.grid_origin) == 3 and len(c.gravity) == 3, "grid_origin and gravity must have three components"),
    Rule(finite_vectors("grid_origin", "gravity"), "grid_origin and gravity must be finite"),
    # the elastic CFL of the undeformed material: a start-up check, not a bound under large deformation
    Rule(lambda c, _: math.isfinite(c.cfl_limit) and 0.0 < c.cfl_limit <= 0.4, "cfl_limit must be in (0, 0.4]"),
    Rule(lambda c, _: re.fullmatch(r"auto|cpu|cuda(?::\d+)?", c.device_name) is not None,
         "device_name must be auto, cpu, cuda, or cuda:<index>"),
    Rule(lambda c, material: not c.time_step * material.wave_speed / c.grid_spacing > c.cfl_limit,
         lambda c: f"initial elastic CFL > {c.cfl_limit}; reduce time_step or stiffness"),
)


@dataclass(frozen=True)
class DingConfig:
    """Ding-2019 viscoplastic porous dough (reference ``config.py:109-145``)."""

    yield_stress: float = 2000.0
    viscosity: float = 1.0e10
    porosity: float = 0.0
    temperature_enabled: bool = False
    temperature_knots_k: tuple = (278.15, 293.15, 313.15, 353.15)
    yield_factors: tuple = (1.5, 1.0, 0.2, 2.0)
    viscosity_factors: tuple = (1.5, 1.0, 0.2, 0.05)
    gas_enabled: bool = False
    gas_initial_pressure_pa: float = 101325.0
    gas_ambient_pressure_pa: float = 101325.0
    gas_production_rate: float = 0.001
    gas_reaction_temperature_k: float = 333.15
    gas_reaction_width_k: float = 15.0

    def Validate(self):
        check_rules(self, DING_RULES)


_TABLES = ("temperature_knots_k", "yield_factors", "viscosity_factors")


def _table_positive(name):
    return Rule(lambda c, _: all(math.isfinite(v) and v > 0 for v in getattr(c, name)),
                f"Ding {name} must contain positive finite values")


def _yield_dips_at_third_knot(c, _):
    y = c.yield_factors
    return y[0] >= y[1] and y[1] >= y[2] and y[2] <= y[3]


DING_RULES = (
    Rule(lambda c, _: positive("yield_stress", "viscosity")(c, _) and math.isfinite(c.porosity) and 0 <= c.porosity < 1,
         "Ding requires yield_stress > 0, viscosity > 0, and 0 <= porosity < 1"),
    Rule(lambda c, _: (nonnegative("gas_initial_pressure_pa", "gas_ambient_pressure_pa", "gas_production_rate")(c, _)
                       and positive("gas_reaction_temperature_k", "gas_reaction_width_k")(c, _)),
         "Gas pressures/rate must be nonnegative; reaction temperatures/width must be positive and finite"),
    Rule(lambda c, _: not c.gas_enabled or c.porosity > 0,
         "Gas mode requires positive initial porosity; this stage has no pore nucleation model"),
    *(_table_positive(name) for name in _TABLES),
    Rule(lambda c, _: all(len(getattr(c, name)) == 4 for name in _TABLES),
         "Ding temperature_knots_k, yield_factors and viscosity_factors require four entries"),
    Rule(lambda c, _: all(a < b for a, b in zip(c.temperature_knots_k, c.temperature_knots_k[1:])),
         "Ding temperature_knots_k must strictly increase"),
    Rule(_yield_dips_at_third_knot, "Ding yield_factors must decrease to the third knot then increase"),
)


@dataclass(frozen=True)
class TransportConfig:
    """Heat, water and settling parameters (reference ``config.py:148-196``)."""

    settling_duration_s: float = 0.0
    settling_auto: bool = False
    settling_min_duration_s: float = 0.1
    settling_hold_duration_s: float = 0.1
    settling_speed_tolerance: float = 0.001
    settling_rate_tolerance: float = 0.001
    enabled: bool = False
    water_enabled: bool = True
    initial_temperature_k: float = 293.15
    oven_temperature_k: float = 453.15
    conductivity: float = 0.4
    specific_heat: float = 2800.0
    heat_transfer: float = 20.0
    tray_temperature_k: float = 453.15
    tray_heat_transfer: float = 100.0
    contact_distance: float = 0.005
    initial_saturation: float = 0.4
    water_density: float = 1000.0
    water_diffusivity: float = 1.0e-10
    surface_water_loss: float = 1.0e-7
    water_stress: float = 10000.0
    latent_heat: float = 0.0

    def Validate(self, material):
        """The settling rules always; the transport rules when transport is enabled."""
        check_rules(self, SETTLING_RULES)
        if self.enabled:
            check_rules(self, TRANSPORT_RULES, material)


def _settling_window_ok(c, _):
    if not c.settling_auto:
        return True
    return (nonnegative("settling_min_duration_s")(c, _)
            and positive("settling_hold_duration_s", "settling_speed_tolerance", "settling_rate_tolerance")(c, _))


SETTLING_RULES = (
    Rule(nonnegative("settling_duration_s"), "settling_duration_s must be finite and nonnegative"),
    Rule(lambda c, _: c.enabled or not c.settling_duration_s > 0,
         "Settling requires transport.enabled=true for local constant-temperature initialization"),
    Rule(_settling_window_ok, "Automatic settling requires nonnegative minimum time and positive hold time/tolerances"),
    Rule(lambda c, _: not c.settling_auto or c.settling_duration_s >= c.settling_min_duration_s + c.settling_hold_duration_s,
         "Maximum settling duration must cover minimum plus hold duration"),
)

TRANSPORT_RULES = (
    Rule(lambda c, material: material.model == "ding2019", "Transport requires the Ding material"),
    Rule(positive("initial_temperature_k", "oven_temperature_k", "tray_temperature_k", "specific_heat", "water_density"),
         "Transport temperatures, heat capacity and water density must be positive"),
    Rule(nonnegative("conductivity", "heat_transfer", "tray_heat_transfer", "contact_distance", "water_diffusivity",
                     "surface_water_loss", "water_stress", "latent_heat"),
         "Transport coefficients must be nonnegative"),
    Rule(lambda c, material: 0 <= c.initial_saturation < 1 and material.density > c.water_density * c.initial_saturation,
         "Initial saturation must be in [0,1), leaving positive dry mass"),
)
