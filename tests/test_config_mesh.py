"""Scene-parameter and mesh validation: the reference's rules and texts (config.py:10-196,
geometry/mesh_validation.py:18-47, tests/test_mesh_processing.py:8-31).

Each case breaks one rule of a valid configuration and expects the reference's message; where several
rules break, the first in the reference's order is reported.
"""
from dataclasses import replace

import numpy as np
import pytest

from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
from baking_amd.geometry.mesh_loader import CreateDemoMesh
from baking_amd.geometry.mesh_validation import ComputeSignedVolumes, ValidateTetMesh

NAN = float("nan")


def test_defaults_validate():
    MaterialConfig().Validate()
    SimulationConfig().Validate(MaterialConfig())
    DingConfig().Validate()
    TransportConfig().Validate(MaterialConfig())
    TransportConfig(enabled=True).Validate(MaterialConfig(model="ding2019"))


@pytest.mark.parametrize("changes,message", [
    (dict(density=0.0), "density must be positive and finite"),
    (dict(density=NAN), "density must be positive and finite"),
    (dict(young_modulus=-1.0), "young_modulus must be positive and finite"),
    (dict(model="bread"), "material model must be neo_hookean or ding2019"),
    (dict(poisson_ratio=0.5), r"poisson_ratio must be in \(-1, 0.5\)"),
    (dict(density=-1.0, model="bread"), "density must be positive"),
])
def test_material_rules(changes, message):
    with pytest.raises(ValueError, match=message):
        replace(MaterialConfig(), **changes).Validate()


@pytest.mark.parametrize("changes,message", [
    (dict(grid_resolution=7), "grid_resolution >= 8 and boundary_cells >= 2 are required"),
    (dict(boundary_cells=1), "grid_resolution >= 8 and boundary_cells >= 2 are required"),
    (dict(grid_resolution=8, boundary_cells=3), "boundary layers leave no usable interior"),
    (dict(grid_spacing=0.0), "grid_spacing must be positive and finite"),
    (dict(time_step=float("inf")), "time_step must be positive and finite"),
    (dict(steps_per_frame=0), "steps_per_frame must be positive"),
    (dict(gravity=(0.0, -9.8)), "grid_origin and gravity must have three components"),
    (dict(grid_origin=(0.0, NAN, 0.0)), "grid_origin and gravity must be finite"),
    (dict(cfl_limit=0.5), r"cfl_limit must be in \(0, 0.4\]"),
    (dict(device_name="gpu"), "device_name must be auto, cpu, cuda, or cuda:<index>"),
    (dict(time_step=1.0), "initial elastic CFL > 0.4; reduce time_step or stiffness"),
])
def test_simulation_rules(changes, message):
    with pytest.raises(ValueError, match=message):
        replace(SimulationConfig(), **changes).Validate(MaterialConfig())
    replace(SimulationConfig(), device_name="cuda:3").Validate(MaterialConfig())


def test_simulation_checks_its_material_first():
    with pytest.raises(ValueError, match="density"):
        replace(SimulationConfig(), grid_resolution=2).Validate(MaterialConfig(density=-1.0))


@pytest.mark.parametrize("changes,message", [
    (dict(yield_stress=0.0), "Ding requires yield_stress > 0, viscosity > 0, and 0 <= porosity < 1"),
    (dict(porosity=1.0), "Ding requires yield_stress > 0"),
    (dict(viscosity=NAN), "Ding requires yield_stress > 0"),
    (dict(gas_production_rate=-1.0), "Gas pressures/rate must be nonnegative"),
    (dict(gas_reaction_width_k=0.0), "reaction temperatures/width must be positive and finite"),
    (dict(gas_enabled=True), "Gas mode requires positive initial porosity"),
    (dict(yield_factors=(1.5, 1.0, 0.0, 2.0)), "Ding yield_factors must contain positive finite values"),
    (dict(viscosity_factors=(1.0, 1.0, 1.0)), "require four entries"),
    (dict(temperature_knots_k=(278.15, 278.15, 313.15, 353.15)), "temperature_knots_k must strictly increase"),
    (dict(yield_factors=(1.0, 1.5, 0.2, 2.0)), "Ding yield_factors must decrease to the third knot then increase"),
])
def test_ding_rules(changes, message):
    with pytest.raises(ValueError, match=message):
        replace(DingConfig(), **changes).Validate()
    replace(DingConfig(), gas_enabled=True, porosity=0.05).Validate()


@pytest.mark.parametrize("changes,message", [
    (dict(settling_duration_s=-1.0), "settling_duration_s must be finite and nonnegative"),
    (dict(settling_duration_s=0.1), "Settling requires transport.enabled=true"),
    (dict(enabled=True, settling_duration_s=1.0, settling_auto=True, settling_hold_duration_s=0.0),
     "Automatic settling requires nonnegative minimum time and positive hold time/tolerances"),
    (dict(enabled=True, settling_duration_s=0.15, settling_auto=True), "Maximum settling duration must cover"),
    (dict(enabled=True, oven_temperature_k=0.0), "Transport temperatures, heat capacity and water density must be positive"),
    (dict(enabled=True, latent_heat=-1.0), "Transport coefficients must be nonnegative"),
    (dict(enabled=True, initial_saturation=1.0), r"Initial saturation must be in \[0,1\), leaving positive dry mass"),
    (dict(enabled=True, initial_saturation=0.9, water_density=2000.0), "leaving positive dry mass"),
])
def test_transport_rules(changes, message):
    with pytest.raises(ValueError, match=message):
        replace(TransportConfig(), **changes).Validate(MaterialConfig(model="ding2019", density=1500.0))


def test_transport_needs_ding_only_when_enabled():
    TransportConfig().Validate(MaterialConfig())
    with pytest.raises(ValueError, match="Transport requires the Ding material"):
        TransportConfig(enabled=True).Validate(MaterialConfig())


def test_mesh_volumes_orientation_and_rejections():
    """tests/test_mesh_processing.py:8-31: the demo cube's volume and positive orientation, repeated indices and
    degenerate tets rejected; volumes equal the triple product e1 . (e2 x e3) / 6 computed independently, and a
    mesh with every tet inverted comes back positively oriented with the same volumes."""
    mesh = CreateDemoMesh(2)
    np.testing.assert_allclose(mesh.reference_volumes.sum(), 0.3 ** 3)
    assert np.all(ComputeSignedVolumes(mesh.vertex_positions, mesh.tet_indices) > 0.0)
    c = mesh.vertex_positions[mesh.tet_indices]
    triple = np.einsum("ij,ij->i", c[:, 1] - c[:, 0], np.cross(c[:, 2] - c[:, 0], c[:, 3] - c[:, 0])) / 6.0
    np.testing.assert_allclose(mesh.reference_volumes, triple, rtol=1e-13)
    flipped = mesh.tet_indices[:, [0, 2, 1, 3]]
    again = ValidateTetMesh(mesh.vertex_positions, flipped)
    np.testing.assert_allclose(again.reference_volumes, mesh.reference_volumes, rtol=1e-13)
    assert np.all(ComputeSignedVolumes(again.vertex_positions, again.tet_indices) > 0.0)
    assert again.tet_indices.flags["C_CONTIGUOUS"] and again.vertex_positions.flags["C_CONTIGUOUS"]
    small = CreateDemoMesh(1)
    with pytest.raises(ValueError, match="repeated"):
        ValidateTetMesh(small.vertex_positions, np.array([[0, 0, 1, 2]]))
    with pytest.raises(ValueError, match="degenerate"):
        ValidateTetMesh(np.zeros((4, 3)), np.array([[0, 1, 2, 3]]))
    with pytest.raises(ValueError, match="duplicate tetrahedra"):
        ValidateTetMesh(small.vertex_positions, np.vstack([small.tet_indices[:1], small.tet_indices[:1, [1, 0, 2, 3]]]))
    with pytest.raises(ValueError, match="outside vertex array"):
        ValidateTetMesh(small.vertex_positions, np.array([[0, 1, 2, 10 ** 6]]))
    with pytest.raises(ValueError, match="integers"):
        ValidateTetMesh(small.vertex_positions, np.array([[0.0, 1.0, 2.0, 3.0]]))
    with pytest.raises(ValueError, match="shape"):
        ValidateTetMesh(small.vertex_positions[:, :2], small.tet_indices)
