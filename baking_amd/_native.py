"""ctypes binding of ``libbaking_hip.so`` (the C ABI declared in ``include/baking_abi.h``).

The library is required: there is no CPU fallback.  Importing this module on a
machine without the built library raises ``ImportError``; calling into it
without a GPU raises ``RuntimeError`` from ``require_gpu``.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint32, c_void_p
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libbaking_hip.so"

F32P = POINTER(c_float)
F64P = POINTER(c_double)
I32P = POINTER(c_int32)


class BkParams(ctypes.Structure):
    _fields_ = [
        ("origin", c_float * 3), ("dx", c_float), ("res", c_int32), ("boundary_cells", c_int32),
        ("gravity", c_float * 3), ("dt", c_float), ("cfl_limit", c_float),
        ("model", c_int32), ("mu", c_float), ("lam", c_float), ("density", c_float),
        ("yield_stress", c_float), ("n0", c_float), ("temperature_enabled", c_int32), ("gas_enabled", c_int32),
        ("knots", c_float * 4), ("yield_factors", c_float * 4), ("viscosity_factors", c_float * 4),
        ("p_ambient", c_float), ("gas_peak", c_float), ("gas_width", c_float),
        ("transport_enabled", c_int32), ("water_enabled", c_int32), ("settling", c_int32),
        ("T_oven", c_float), ("T_tray", c_float), ("tray_h", c_float), ("loss_coef", c_float), ("rho_w", c_float),
        ("latent", c_float), ("kappa", c_float), ("diffusivity", c_float), ("s0", c_float), ("water_stress", c_float),
        ("contact_distance", c_float),
        ("n", c_int32), ("n_vertices", c_int32), ("n_faces", c_int32), ("n_slots", c_int32),
        ("return_solver", c_int32), ("surrogate", c_int32), ("sticky_floor", c_int32),
    ]


STATE_FIELDS = ["x", "v", "F", "C", "Fe", "Fp", "tau", "residual", "porosity", "T_mat", "gelled", "yield_stress", "viscosity",
                "porosity_residual", "density", "moles", "pressure", "T", "W", "s", "losses", "exchanges", "masses"]


class BkState(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in STATE_FIELDS]


SCENE_FIELDS = ["V0", "m0", "capacities", "dry", "visc_param", "prod_param", "heat_param", "ref_vertices", "bind_offsets",
                "bind_particles", "bind_weights", "bind_ref_offsets", "faces", "slot_face_offsets", "slot_faces", "slot_owner",
                "particle_slot", "dough_of"]


class BkScene(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in SCENE_FIELDS]


class BkDough(ctypes.Structure):
    _fields_ = [("params", BkParams), ("particle_offset", c_int32), ("block_offset", c_int32)]


class BkTuning(ctypes.Structure):
    _fields_ = [("tile_chunk", c_int32), ("particle_block", c_int32), ("grid_cap", c_int32), ("deferred_wgs", c_int32)]


GRID_FIELDS = ["m", "p", "v", "V", "C", "E", "W", "H", "L", "X", "dT", "dW"]


class BkGrid(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in GRID_FIELDS]


ADJOINT_FIELDS = ["x", "v", "C", "F", "Fe", "tau", "porosity", "moles", "T", "W", "losses", "exchanges", "rev_offsets", "rev_bindings",
                  "rev_vertices", "anchor_coef"]


class BkAdjoint(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in ADJOINT_FIELDS]


class BkSettling(ctypes.Structure):
    """Settling controller record of one dough (``bk_settling``)."""
    _fields_ = [("auto_mode", c_int32), ("settling_steps", c_int32), ("time_step", c_double), ("min_duration_s", c_double),
                ("hold_duration_s", c_double), ("speed_tolerance", c_double), ("rate_tolerance", c_double),
                ("phase", c_int32), ("reason", c_int32), ("step_count", c_int64), ("reference_step", c_int64),
                ("quiet_steps", c_int64), ("baking_start_step", c_int64), ("plastic_rate", c_double),
                ("diagnostics", c_float * 3), ("ref_slot", c_int32), ("copy_pending", c_int32), ("reserved", c_int32)]


SETTLING_REASONS = ("pending", "fixed_duration", "quasi_static", "timeout")

# out[] slots of bk_check_state
CHECK_FIELDS = ("nonfinite_state", "minimum_j", "maximum_j", "outside_stencil", "maximum_speed", "gas_wave", "water_wave_sq",
                "nonfinite_ding", "maximum_residual", "nonfinite_vertices", "inverted_tets", "particle_volume", "total_mass",
                "grid_mass", "kinetic_energy")


# Exported symbols and their argument types; the ABI test checks this list
# against the header and the library.
P = c_void_p
F = c_float
I = c_int32
L = c_int64
SIGNATURES = {
    "bk_abi_version": ([], c_int),
    "bk_last_error": ([], c_char_p),
    "bk_device_info": ([c_char_p, c_int, POINTER(c_int), POINTER(c_int)], c_int),
    "bk_workspace_create": ([I, I, I, I], c_void_p),
    "bk_workspace_create_batch": ([I, I, I, I, I], c_void_p),
    "bk_workspace_destroy": ([P], None),
    "bk_workspace_set_doughs": ([P, POINTER(BkDough), I], c_int),
    "bk_workspace_set_tuning": ([P, POINTER(BkTuning)], c_int),
    "bk_workspace_get_tuning": ([P, POINTER(BkTuning)], c_int),
    "bk_workspace_flags": ([P], c_void_p),
    "bk_workspace_vertices": ([P], c_void_p),
    "bk_workspace_counts": ([P, P, P], c_int),
    "bk_workspace_sort_order": ([P, P, P, P, I, P], c_int),
    "bk_workspace_read_flags": ([P, P, P], c_int),
    "bk_workspace_write_flags": ([P, P, P], c_int),
    "bk_workspace_read_flags_all": ([P, P, I, P], c_int),
    "bk_workspace_write_flags_all": ([P, P, I, P], c_int),
    "bk_workspace_set_freezing": ([P, I], c_int),
    "bk_workspace_freeze_modes": ([P, P, I, I, P], c_int),
    "bk_cell_keys": ([P, I, P, F, I, I, P, P, P], c_int),
    "bk_sort_pairs": ([P, P, P, P, P, I, I, P], c_int),
    "bk_sort_pairs_noalloc": ([P, P, P, P, I, I, P], c_int),
    "bk_sort_pairs_noalloc_capacity": ([], c_int32),
    "bk_noalloc_last_error": ([], c_char_p),
    "bk_substep": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), P, P], c_int),
    "bk_substep_graph_create": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), P],
                                c_void_p),
    "bk_graph_launch": ([P, P], c_int),
    "bk_batch_substep": ([P, POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), P], c_int),
    "bk_batch_graph_create": ([P, POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid)], c_void_p),
    "bk_substep_graph_create_pipelined": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid),
                                           P], c_void_p),
    "bk_batch_graph_create_pipelined": ([P, POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid)], c_void_p),
    "bk_substep_graph_create_steps": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), P,
                                       I, I], c_void_p),
    "bk_batch_graph_create_steps": ([P, POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), I, I], c_void_p),
    "bk_substep_finish": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), P, P], c_int),
    "bk_batch_finish": ([P, POINTER(BkScene), POINTER(BkState), POINTER(BkState), P], c_int),
    "bk_graph_destroy": ([P], None),
    "bk_profile_begin": ([P, I], c_int),
    "bk_profile_end": ([P, P, P], c_int),
    "bk_profile_stage_count": ([], c_int32),
    "bk_profile_stage_name": ([I], c_char_p),
    "bk_measure": ([POINTER(BkScene), POINTER(BkState), I, I, P, P], c_int),
    "bk_check_scratch_size": ([I, L, I, I], c_int64),
    "bk_check_state": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), P, P, L, P, I, P, I, P, P, P], c_int),
    "bk_grid_untile": ([P, P, I, I, P], c_int),
    "bk_grid_untile_f64": ([P, P, I, P], c_int),
    "bk_ParticleToGrid": ([P, P, P, P, P, P, P, P, P, F, I, F, P, I, P], c_int),
    "bk_UpdateGrid": ([P, P, P, I, I, P, F, L, P], c_int),
    "bk_GridToParticle": ([P, P, P, P, P, P, P, P, F, I, F, P, I, P], c_int),
    "bk_ComputeElasticStress": ([P, P, F, F, P, I, P], c_int),
    "bk_UpdateDing": ([P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I, F, F, F, F, I, P, P, P, F, P, I, P], c_int),
    "bk_CombineGasStress": ([P, P, P, P, P, P, P, F, F, P, F, F, F, F, I, P], c_int),
    "bk_RefreshGas": ([P, P, P, P, P, P, I, P], c_int),
    "bk_SurfaceAreas": ([P, P, P, P, P, F, F, I, P], c_int),
    "bk_ScalarToGrid": ([P, P, P, P, P, P, P, P, P, F, F, F, F, F, F, I, P, P, P, P, P, P, P, P, P, P, F, I, I, P], c_int),
    "bk_GridFields": ([P, P, P, P, P, P, F, L, P], c_int),
    "bk_DiffuseGrid": ([P, P, P, P, P, P, P, P, P, P, P, F, F, F, I, F, I, F, L, P], c_int),
    "bk_ScalarToParticles": ([P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, F, P, F, I, F, I, P], c_int),
    "bk_LocalGas": ([P, P, P, P, P, P, P, P, F, F, F, F, I, I, P], c_int),
    "bk_WaterStress": ([P, P, P, P, F, F, I, P], c_int),
    "bk_Saturations": ([P, P, P, P, F, I, P], c_int),
    "bk_ReconstructMeshVertices": ([P, P, P, P, P, P, P, P, I, P], c_int),
    "bk_gas_pressure": ([P, P, P, P, P, I, P], c_int),
    "bk_porosity_coefficient": ([P, P, I, P], c_int),
    "bk_solve_porosity": ([P, P, P, P, I, I, P], c_int),
    "bk_mlp_theta_count": ([], c_int32),
    "bk_mlp_blob_bytes": ([], c_int32),
    "bk_mlp_grad_chunk": ([], c_int32),
    "bk_mlp_infer": ([P, P, P, I, P], c_int),
    "bk_mlp_forward": ([P, P, P, I, P, P], c_int),
    "bk_mlp_gradient": ([P, P, P, I, P, P, P, P, P, P, P, P], c_int),
    "bk_adam_step": ([P, P, P, P, I, F, F, F, F, I, P], c_int),
    "bk_workspace_set_surrogate": ([P, P], c_int),
    "bk_workspace_set_samples": ([P, P, I], c_int),
    "bk_workspace_stats": ([P, P, I, P], c_int),
    "bk_workspace_settling_set": ([P, I, I, POINTER(BkSettling), P], c_int),
    "bk_workspace_settling_get": ([P, I, I, POINTER(BkSettling), P], c_int),
    "bk_workspace_settling_reference": ([P, P], c_int),
    "bk_SettlingMetrics": ([P, P, P, P, P, F, I, P], c_int),
    "bk_adjoint_scratch_size": ([POINTER(BkParams)], c_int64),
    "bk_adjoint_substep": ([POINTER(BkParams), POINTER(BkScene), POINTER(BkState), POINTER(BkState), POINTER(BkGrid), P,
                            POINTER(BkAdjoint), P, P, P], c_int),
    "bk_UpdateDing_adjoint": ([P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, F, F, F, F, I, P, P, P, F, I, P], c_int),
    "bk_overstress_adjoint": ([P, P, P, I, P], c_int),
    "bk_solve_porosity_adjoint": ([P, P, P, P, I, P], c_int),
}

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load():
    """Load the HIP library (no GPU needed to load; symbols only)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(f"{_LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(str(_LIB_PATH))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("baking_amd runs its substep only on a ROCm GPU (MI355X); no CPU fallback exists")


def check(rc: int, what: str):
    if rc != 0:
        msg = load().bk_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("baking_amd kernels need contiguous tensors")
    return c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def host_f32(values):
    arr = (c_float * len(values))(*[float(v) for v in values])
    return ctypes.cast(arr, c_void_p), arr
