"""ctypes binding of the in-tree native library ``libems_gpu.so``.

The library holds the HIP engine (include/ems_gpu.h), the C++ host side
(include/ems_host.h) and the early-stop network (include/ems_learn.h).
There is no fallback: if the library is missing the import fails with the
command that builds it.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMS_AMD_FP_CONTRACT=fast selects the build whose step kernels use fused
# multiply-adds (results within rounding of the reference engine instead of
# bit-identical).
FP_CONTRACT = os.environ.get("EMS_AMD_FP_CONTRACT", "off")
if FP_CONTRACT not in ("off", "fast"):
    raise ImportError(f"EMS_AMD_FP_CONTRACT must be 'off' or 'fast', not {FP_CONTRACT!r}")
LIB_PATH = os.path.join(_HERE, "libems_gpu.so" if FP_CONTRACT == "off" else "libems_gpu_fma.so")

EMS_OK = 0
EMS_E_INVALID = -1
EMS_E_UNSUPPORTED = -2
EMS_E_IO = -3
EMS_E_DEVICE = -4
EMS_PROBE_F32_SUM = 4  # or-ed into a probe's quantity


class EmsError(RuntimeError):
    """An ``ems::Error`` from the native side (src/error.rs variants)."""

    KINDS = {EMS_E_INVALID: "Invalid", EMS_E_UNSUPPORTED: "Unsupported",
             EMS_E_IO: "Io", EMS_E_DEVICE: "Device"}

    def __init__(self, code: int, msg: str):
        self.code = code
        self.kind = self.KINDS.get(code, "Unknown")
        super().__init__(f"{self.kind}: {msg}")


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (hipcc --offload-arch=gfx950); there is no CPU fallback")
    return C.CDLL(LIB_PATH)


lib = _load()

_d = C.c_double
_i = C.c_int
_u64 = C.c_uint64
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)


class GridDesc(C.Structure):
    _fields_ = [("nodes", C.c_int32 * 3), ("dt", C.c_double), ("curl_order", C.c_int32)]


class CpmlSlab(C.Structure):
    _fields_ = [("e_side", C.c_int32), ("comp", C.c_int32), ("normal", C.c_int32),
                ("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3),
                ("b", _fp), ("a", _fp), ("k", _fp)]


class CpmlParams(C.Structure):
    _fields_ = [("cells", C.c_int32), ("grading_order", C.c_double),
                ("sigma_factor", C.c_double), ("kappa_max", C.c_double),
                ("alpha_max", C.c_double), ("convolution", C.c_int32)]


class RunLimitsC(C.Structure):
    _fields_ = [("max_timesteps", C.c_uint64), ("has_criteria", C.c_int32),
                ("energy_ratio", C.c_double), ("check_interval", C.c_uint64),
                ("has_excitation_end", C.c_int32), ("excitation_end", C.c_double),
                ("has_settle", C.c_int32), ("settle_frequency", C.c_double),
                ("settle_tolerance", C.c_double), ("settle_window", C.c_double),
                ("settle_max_extension", C.c_double), ("check_after_interval", C.c_int32),
                ("has_extra_check", C.c_int32), ("extra_check", C.c_uint64),
                ("energy_measure", C.c_int32)]


class SolidC(C.Structure):
    """ems_solid (include/ems_host.h)."""
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("p", C.c_double * 6)]


class PoleEntryC(C.Structure):
    """ems_pole_entry (include/ems_host.h)."""
    _fields_ = [("field", C.c_int32), ("axis", C.c_int32), ("node", C.c_int32 * 3),
                ("kind", C.c_int32), ("p", C.c_double * 3)]


class PolePlanC(C.Structure):
    """ems_pole_plan (include/ems_host.h)."""
    _fields_ = [("n_edges", C.c_int32 * 2), ("axis", C.POINTER(C.c_int32) * 2),
                ("nodes", C.POINTER(C.c_int32) * 2), ("sum_inv_z", C.POINTER(C.c_double) * 2),
                ("pole_end", C.POINTER(C.c_uint32) * 2), ("n_poles", C.c_uint64 * 2),
                ("cidx", C.POINTER(C.c_uint32) * 2), ("n_consts", C.c_int32),
                ("a", C.POINTER(C.c_double)), ("inv_c", C.POINTER(C.c_double)),
                ("inv_z", C.POINTER(C.c_double)), ("half_l", C.POINTER(C.c_double)),
                ("kind", C.POINTER(C.c_int32)), ("skipped", C.c_int32)]


class LumpedElementC(C.Structure):
    """ems_lumped_element (include/ems_host.h)."""
    _fields_ = [("axis", C.c_int32), ("n_edges", C.c_int32), ("nodes", C.POINTER(C.c_int32)),
                ("series", C.c_int32), ("columns", C.c_int32), ("r", C.c_double),
                ("l", C.c_double), ("c", C.c_double), ("edge_r", C.POINTER(C.c_double)),
                ("edge_l", C.POINTER(C.c_double)), ("edge_c", C.POINTER(C.c_double)),
                ("topology", C.c_int32), ("explicit_inductor", C.c_int32)]


class LumpedInstallC(C.Structure):
    """ems_lumped_install (include/ems_host.h)."""
    _fields_ = [("cb_e", C.POINTER(C.c_double)), ("g_src", C.POINTER(C.c_double)),
                ("rec_g", C.POINTER(C.c_double)), ("rec_c", C.POINTER(C.c_double)),
                ("two_l_dt", C.POINTER(C.c_double)), ("inv_c", C.POINTER(C.c_double)),
                ("z", C.POINTER(C.c_double)), ("branch", C.c_int32), ("src_in_branch", C.c_int32)]


class RunOutcomeC(C.Structure):
    _fields_ = [("reason", C.c_int32), ("ratio_db", C.c_double), ("steps", C.c_uint64),
                ("timestep", C.c_uint64), ("time", C.c_double), ("energy", C.c_double),
                ("max_energy", C.c_double), ("elapsed_s", C.c_double),
                ("has_criterion_steps", C.c_int32), ("criterion_steps", C.c_uint64)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


# include/ems_gpu.h
ems_gpu_create = _sig("ems_gpu_create", _i, C.POINTER(GridDesc), C.POINTER(_p))
ems_gpu_destroy = _sig("ems_gpu_destroy", None, _p)
ems_gpu_last_error = _sig("ems_gpu_last_error", C.c_char_p, _p)
ems_gpu_set_lane = _sig("ems_gpu_set_lane", _i, _p, _i, _fp, _u64, _u32p)
ems_gpu_add_cpml = _sig("ems_gpu_add_cpml", _i, _p, C.POINTER(CpmlSlab), _i)
ems_gpu_add_mur = _sig("ems_gpu_add_mur", _i, _p, _p, C.POINTER(_i))
ems_gpu_add_upml = _sig("ems_gpu_add_upml", _i, _p, _p, C.POINTER(_i))
ems_gpu_upml_state = _sig("ems_gpu_upml_state", _i, _p, _i, _i, _fp, _fp)
ems_gpu_add_source = _sig("ems_gpu_add_source", _i, _p, _i, _i, _ip, _ip, _fp, _fp, _u64,
                          C.POINTER(_i))
ems_gpu_add_probe = _sig("ems_gpu_add_probe", _i, _p, _i, _i, _ip, _ip, _dp, _u64,
                         C.POINTER(_i))
ems_gpu_add_lumped = _sig("ems_gpu_add_lumped", _i, _p, _p, C.POINTER(_i))
ems_gpu_lumped_table = _sig("ems_gpu_lumped_table", _i, _p, _i, _i, _dp, _u64)
ems_gpu_lumped_values = _sig("ems_gpu_lumped_values", _i, _p, _i, _i, _dp, _u64, C.POINTER(_u64),
                             C.POINTER(_u64))
ems_gpu_lumped_branch = _sig("ems_gpu_lumped_branch", _i, _p, _i, _dp, _dp, _dp)
ems_gpu_add_dft = _sig("ems_gpu_add_dft", _i, _p, _i, _i, _dp, _i, _d, _d, C.POINTER(_i))
ems_gpu_dft_update = _sig("ems_gpu_dft_update", _i, _p, _i)
ems_gpu_dft_result = _sig("ems_gpu_dft_result", _i, _p, _i, _dp, C.POINTER(_u64))
ems_gpu_add_dispersive = _sig("ems_gpu_add_dispersive", _i, _p, _p, _p, _p, C.POINTER(_i))
ems_gpu_dispersive_state = _sig("ems_gpu_dispersive_state", _i, _p, _i, _i, _dp, _dp)
ems_gpu_add_planewave = _sig("ems_gpu_add_planewave", _i, _p, _p, C.POINTER(_i))
ems_gpu_planewave_table = _sig("ems_gpu_planewave_table", _i, _p, _i, _dp, _u64)
ems_gpu_planewave_aux = _sig("ems_gpu_planewave_aux", _i, _p, _i, _dp, _dp)
ems_gpu_add_merged = _sig("ems_gpu_add_merged", _i, _p, _p, C.POINTER(_i))
ems_gpu_merged_state = _sig("ems_gpu_merged_state", _i, _p, _i, _dp, _dp)
ems_gpu_batch_create = _sig("ems_gpu_batch_create", _i, C.POINTER(_p), _i, C.POINTER(_p))
ems_gpu_energy_many = _sig("ems_gpu_energy_many", _i, C.POINTER(_p), _i,
                           C.POINTER(C.c_uint8), _dp)
ems_gpu_batch_destroy = _sig("ems_gpu_batch_destroy", None, _p)
ems_gpu_batch_last_error = _sig("ems_gpu_batch_last_error", C.c_char_p, _p)
ems_gpu_batch_set_schedule = _sig("ems_gpu_batch_set_schedule", _i, _p, _i, _i, _i)
ems_gpu_batch_set_graphs = _sig("ems_gpu_batch_set_graphs", _i, _p, _i)
ems_gpu_batch_run = _sig("ems_gpu_batch_run", _i, _p, _u8p, _u64)

# include/ems_learn.h
ems_cn_create = _sig("ems_cn_create", _i, _i, _i, _i, _i, _ip, _u64, C.POINTER(_p))
ems_cn_destroy = _sig("ems_cn_destroy", None, _p)
ems_cn_last_error = _sig("ems_cn_last_error", C.c_char_p, _p)
ems_cn_param_count = _sig("ems_cn_param_count", _u64, _p)
ems_cn_set_params = _sig("ems_cn_set_params", _i, _p, _fp, _u64)
ems_cn_get_params = _sig("ems_cn_get_params", _i, _p, _fp, _u64)
ems_cn_set_data = _sig("ems_cn_set_data", _i, _p, _fp, _fp, _i, _i)
ems_cn_set_spectral = _sig("ems_cn_set_spectral", _i, _p, _fp, _i, _i, C.c_float)
ems_cn_train_batch = _sig("ems_cn_train_batch", _i, _p, _ip, _i, _i, C.c_float, C.c_float, _fp,
                          _fp)
ems_cn_forward = _sig("ems_cn_forward", _i, _p, _fp, _i, _i, _fp)
ems_cn_loss_grad = _sig("ems_cn_loss_grad", _i, _p, _fp, _fp, _i, _i, _i, _fp, _fp)
ems_gpu_run = _sig("ems_gpu_run", _i, _p, _u64)
ems_gpu_step_with_energy = _sig("ems_gpu_step_with_energy", _i, _p, _dp)
ems_gpu_energy = _sig("ems_gpu_energy", _i, _p, _dp)
ems_gpu_energy_noalloc = _sig("ems_gpu_energy_noalloc", _i, _p, _dp)
ems_gpu_energy_in = _sig("ems_gpu_energy_in", _i, _p, _i, _dp)
ems_gpu_timestep = _sig("ems_gpu_timestep", _u64, _p)
ems_gpu_read_component = _sig("ems_gpu_read_component", _i, _p, _i, _i, _fp)
ems_gpu_write_component = _sig("ems_gpu_write_component", _i, _p, _i, _i, _fp)
ems_gpu_probe_values = _sig("ems_gpu_probe_values", _i, _p, _i, _dp, _u64, C.POINTER(_u64))
ems_gpu_probe_times = _sig("ems_gpu_probe_times", _i, _p, _i, _dp, _u64, C.POINTER(_u64))
ems_gpu_dt = _sig("ems_gpu_dt", _d, _p)
ems_gpu_synchronize = _sig("ems_gpu_synchronize", _i, _p)
ems_gpu_time_kernels = _sig("ems_gpu_time_kernels", _i, _p, _u64, _dp)
ems_gpu_traffic_model = _sig("ems_gpu_traffic_model", _i, _p, _dp)
ems_gpu_set_schedule = _sig("ems_gpu_set_schedule", _i, _p, _i, _i, _i)
ems_gpu_set_xcd_group = _sig("ems_gpu_set_xcd_group", _i, _p, _i)
ems_gpu_set_cpml_split = _sig("ems_gpu_set_cpml_split", _i, _p, _i)
ems_gpu_set_tile_classes = _sig("ems_gpu_set_tile_classes", _i, _p, _i)
ems_gpu_set_unfused_prefetch = _sig("ems_gpu_set_unfused_prefetch", _i, _p, _i)
ems_gpu_copy_bandwidth = _sig("ems_gpu_copy_bandwidth", _i, _p, _u64, _i, _dp)

# include/ems_host.h
ems_courant_limit = _sig("ems_courant_limit", _d, _dp, _i, _dp, _i, _dp, _i)
ems_coefficients_from_cells = _sig("ems_coefficients_from_cells", _i, _dp, _i, _dp, _i, _dp, _i,
                                   _u16p, _dp, _i, _u8p, _ip, _d, _i, _i, _fp)
ems_edge_materials_from_cells = _sig("ems_edge_materials_from_cells", _i, _dp, _i, _dp, _i, _dp,
                                     _i, _u16p, _dp, _i, _i, _dp, _dp)
ems_raster_cells = _sig("ems_raster_cells", C.c_int64, _dp, _i, _dp, _i, _dp, _i,
                        C.POINTER(SolidC), C.c_uint16, _u16p)
ems_pool_lane = _sig("ems_pool_lane", _i, _fp, _i, _i, _i, _fp, _u32p, C.POINTER(_u64))
ems_dispersive_fill_cells = _sig("ems_dispersive_fill_cells", _i, _dp, _i, _dp, _i, _dp, _i,
                                 C.POINTER(C.c_int32), _i, C.POINTER(PoleEntryC), _i,
                                 C.POINTER(PoleEntryC), _u64, C.POINTER(_u64))
ems_dispersive_plan = _sig("ems_dispersive_plan", _i, _dp, _i, _dp, _i, _dp, _i, _d, _fp, _i,
                           C.POINTER(PoleEntryC), _i, C.POINTER(PolePlanC))
ems_lumped_plan = _sig("ems_lumped_plan", _i, C.POINTER(C.c_int32), _d,
                       C.POINTER(LumpedElementC), _fp, _fp, C.POINTER(LumpedInstallC))
ems_cpml_coefficients = _sig("ems_cpml_coefficients", None, C.POINTER(CpmlParams), _d, _d, _d,
                             _d, _dp)
ems_cpml_layer_index = _sig("ems_cpml_layer_index", _d, _dp, _i, _dp, _i, _dp, _i, _fp, _d, _i, _i)
ems_cpml_layer = _sig("ems_cpml_layer", _i, _dp, _i, _dp, _i, _dp, _i, _i, C.POINTER(CpmlParams),
                      _d, _d, C.POINTER(CpmlSlab), _fp)
ems_cpml_profile = _sig("ems_cpml_profile", _i, _dp, _i, _dp, _i, _dp, _i, _i,
                        C.POINTER(CpmlParams), _d, _d, _i, _dp, C.POINTER(_i))
ems_gaussian = _sig("ems_gaussian", _d, _d, _d, _d)
ems_gaussian_duration = _sig("ems_gaussian_duration", _d, _d)
ems_gaussian_table = _sig("ems_gaussian_table", None, _d, _d, _d, _d, _u64, _fp)
ems_gaussian_values = _sig("ems_gaussian_values", None, _d, _d, _d, _d, _u64, _dp)
ems_openems_gaussian = _sig("ems_openems_gaussian", _d, _d, _d, _d, _d)
ems_openems_gaussian_duration = _sig("ems_openems_gaussian_duration", _d, _d)
ems_openems_source_table = _sig("ems_openems_source_table", _i, _d, _d, _d, _i, _d, _u64, _fp)
ems_fold_parallel = _sig("ems_fold_parallel", _i, C.c_float, C.c_float, _d, _d, _d,
                         C.POINTER(C.c_float), C.POINTER(C.c_float))
ems_dft = _sig("ems_dft", None, _d, _d, _dp, _u64, _dp, _i, _dp)
ems_dft_phasors = _sig("ems_dft_phasors", None, _dp, _i, _d, _d, _u64, _dp)
ems_dft_steps = _sig("ems_dft_steps", None, _dp, _i, _d, _dp)
ems_conformal_stability = _sig("ems_conformal_stability", _d, _p)
ems_run_until = _sig("ems_run_until", _i, _p, C.POINTER(RunLimitsC), _ip, _i,
                     C.POINTER(RunOutcomeC))
ems_run_until_batch = _sig("ems_run_until_batch", _i, _p, _i, C.POINTER(_p), C.POINTER(RunLimitsC),
                          C.POINTER(RunOutcomeC))
ems_drive_script = _sig("ems_drive_script", _i, _dp, _u64, _dp, _u64, C.POINTER(RunLimitsC),
                        C.POINTER(RunOutcomeC))


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc: int, handle=None) -> None:
    if rc != EMS_OK:
        msg = ems_gpu_last_error(handle).decode() if handle else ""
        raise EmsError(rc, msg)
