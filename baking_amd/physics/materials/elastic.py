"""Compressible Neo-Hookean material (reference ``physics/materials/elastic.py:1-49``)."""
from ...launch import Kernel

ComputeElasticStress = Kernel("bk_ComputeElasticStress", ["f32", "f32", "f", "f", "i32"])


class ElasticMaterial:
    """psi = mu/2 (I1 - 3) - mu log J + lambda/2 log^2 J; transfers consume Kirchhoff stress."""

    def __init__(self, material_config):
        material_config.Validate()
        self.material_config = material_config

    def ComputeStress(self, particle_state, kirchhoff_stresses, error_flags):
        ComputeElasticStress(len(kirchhoff_stresses), particle_state.deformation_gradients, kirchhoff_stresses,
                             self.material_config.shear_modulus, self.material_config.lame_lambda, error_flags)

    def AdvanceState(self, input_state, output_state, time_step):
        output_state.material_state = None
