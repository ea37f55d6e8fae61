"""ctypes binding of libvaikkari_gpe.so (declarations in include/*.h).

The native library is the only implementation: if it is missing, importing
this module raises.  There is no Python fallback for any computation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libvaikkari_gpe.so")

GPE_CLASSES = 12
GPE_FACES = 4

# gpe_status
OK, EINVAL, ENOMEM, EDEVICE, EUNSTABLE, EBUSY, ENOTSUP, EIO, ECANCELED, ETIMEDOUT = range(10)
# gpe_precision
FP64, FP32_BULK = 0, 1
# gpe_mode
REAL_TIME, IMAG_TIME = 0, 1
# gpe_potential_kind
POT_ARRAY, POT_TRAP = 0, 1
# gpe_kernel
KERNEL_AUTO, KERNEL_GATHER, KERNEL_ZMARCH, KERNEL_FUSED = 0, 1, 2, 3
# gpe_init_kind
INIT_HOST, INIT_GAUSSIAN = 0, 1


class GpeConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("precision", C.c_uint32),
        ("block_scale", C.c_double),
        ("dt", C.c_double),
        ("g", C.c_double),
        ("origin", C.c_double * 3),
        ("potential", C.c_uint32),
        ("mode", C.c_uint32),
        ("trap_ratio", C.c_double),
        ("kernel", C.c_uint32),
        ("tile", C.c_uint32),
    ]


class GpeSolverInfo(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("steps_done", C.c_uint64),
        ("kernel", C.c_uint32),
        ("tile", C.c_uint32),
        ("launches_per_step", C.c_uint32),
        ("bytes_per_node_step", C.c_double),
    ]


class GpeReferenceRun(C.Structure):
    _fields_ = [
        ("dims", C.c_uint32 * 3),
        ("steps_per_iteration", C.c_uint32),
        ("origin", C.c_double * 3),
        ("dt", C.c_double),
        ("g", C.c_double),
        ("maxpot", C.c_double),
        ("lapfac_dt", C.c_double),
        ("lapfac0_dt", C.c_double),
    ]


class GpeInstanceConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("mode", C.c_uint32),
        ("block_scale", C.c_double),
        ("dt", C.c_double),
        ("g", C.c_double),
        ("trap_ratio", C.c_double),
        ("origin", C.c_double * 3),
        ("steps", C.c_uint64),
        ("init", C.c_uint32),
        ("zlayers", C.c_uint32),
        ("sigma", C.c_double),
        ("noise", C.c_double),
        ("mu0", C.c_double),
        ("seed", C.c_uint64),
    ]


class GpeInstanceInfo(C.Structure):
    _fields_ = [
        ("steps_done", C.c_uint64),
        ("steps_left", C.c_uint64),
        ("nodes", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("mode", C.c_uint32),
        ("active", C.c_uint32),
    ]


class GpeRequest(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("mode", C.c_uint32), ("steps", C.c_uint64),
        ("h", C.c_double), ("dt", C.c_double), ("g", C.c_double), ("trap_ratio", C.c_double),
        ("sigma", C.c_double), ("noise", C.c_double), ("mu0", C.c_double),
        ("seed", C.c_uint64),
        ("keep_result", C.c_uint32), ("reserved", C.c_uint32),
    ]


class GpeCompletion(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("arrival", C.c_double), ("admitted", C.c_double), ("finished", C.c_double),
        ("node_steps", C.c_double), ("norm", C.c_double), ("energy", C.c_double),
        ("status", C.c_int32), ("reserved", C.c_uint32),
    ]


class GpeServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
        ("queued", C.c_uint64), ("active", C.c_uint64), ("resident_bytes", C.c_uint64),
        ("node_steps_done", C.c_double), ("device_ms", C.c_double), ("ticks", C.c_double),
        ("failed", C.c_uint64), ("cancelled", C.c_uint64), ("timed_out", C.c_uint64),
    ]


class GpeFieldLayout(C.Structure):
    _fields_ = [
        ("elems", C.c_uint64), ("origin", C.c_uint64), ("xp", C.c_uint64), ("sy", C.c_uint64), ("sz", C.c_uint64),
        ("elem_bytes", C.c_uint32), ("reserved", C.c_uint32),
    ]


class GpeRelaxReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("converged", C.c_uint32),
        ("diff", C.c_double),
        ("prob", C.c_double),
        ("mu", C.c_double),
        ("error", C.c_double),
    ]


RZ_POTENTIAL = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_void_p)
RELAX_PROGRESS = C.CFUNCTYPE(None, C.c_uint32, C.c_double, C.c_double, C.c_void_p)


class GpeOutputOptions(C.Structure):
    _fields_ = [
        ("save_picture", C.c_uint32),
        ("save_volume", C.c_uint32),
        ("volume_flags", C.c_uint32),
        ("reserved", C.c_uint32),
        ("prefix", C.c_char_p),
    ]


class GpePrimalConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("precision", C.c_uint32),
        ("a", C.c_double),
        ("dt", C.c_double),
        ("g", C.c_double),
        ("origin", C.c_double * 3),
        ("trap_ratio", C.c_double),
        ("mode", C.c_uint32),
        ("tile", C.c_uint32),
    ]


class GpeError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: status {status} ({detail})")
        self.status = status


_P = C.c_void_p
_D = C.c_double
_U32 = C.c_uint32
_U64 = C.c_uint64
_I = C.c_int
_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)
_I32P = C.POINTER(C.c_int32)

# (name, restype, argtypes) of every exported symbol of include/*.h.
SIGNATURES = [
    ("gpe_create", _I, [C.POINTER(GpeConfig), C.POINTER(_P)]),
    ("gpe_destroy", None, [_P]),
    ("gpe_set_potential", _I, [_P, _DP]),
    ("gpe_set_state", _I, [_P, _DP, _DP]),
    ("gpe_get_state", _I, [_P, _DP, _DP]),
    ("gpe_set_mode", _I, [_P, _U32, _D]),
    ("gpe_init_odd", _I, [_P, _D]),
    ("gpe_init_gaussian", _I, [_P, _D, _D, _U64, _D]),
    ("gpe_step", _I, [_P, _U64]),
    ("gpe_time_steps", _I, [_P, _U64, _DP, C.POINTER(_U64)]),
    ("gpe_observables", _I, [_P, _I, _DP]),
    ("gpe_step_observe", _I, [_P, _U64, _DP]),
    ("gpe_compare_fields", _I, [_P, _I, _P, _I, _DP]),
    ("gpe_get_field_layout", _I, [_U32, _U32, _U32, _U32, C.POINTER(GpeFieldLayout)]),
    ("gpe_get_field_device", _I, [_P, _I, C.POINTER(_P)]),
    ("gpe_observables_noalloc", _I, [C.POINTER(GpeConfig), _P, _P, _P, _P]),
    ("gpe_synchronize", _I, [_P]),
    ("gpe_get_info", _I, [_P, C.POINTER(GpeSolverInfo)]),
    ("gpe_set_variant", _I, [_P, _U32, _U32]),
    ("gpe_variant_tiles", _I, [_U32]),
    ("gpe_variant_shape", None, [_U32, _U32, _I32P]),
    ("gpe_diag_copy_bandwidth", _I, [_U64, _U32, _DP]),
    ("gpe_diag_stream_bandwidth", _I, [_U32, _U64, _U32, _DP]),
    ("gpe_batch_create", _I, [_U32, _U64, C.POINTER(_P)]),
    ("gpe_batch_destroy", None, [_P]),
    ("gpe_batch_add", _I, [_P, C.POINTER(GpeInstanceConfig), C.POINTER(C.c_int32)]),
    ("gpe_batch_instance_bytes", _U64, [_P, C.POINTER(GpeInstanceConfig)]),
    ("gpe_batch_set_state", _I, [_P, C.c_int32, _DP, _DP]),
    ("gpe_batch_get_state", _I, [_P, C.c_int32, _DP, _DP]),
    ("gpe_batch_observables", _I, [_P, C.c_int32, _DP]),
    ("gpe_batch_instance", _I, [_P, C.c_int32, C.POINTER(GpeInstanceInfo)]),
    ("gpe_batch_remove", _I, [_P, C.c_int32]),
    ("gpe_batch_run", _I, [_P, _U32, _U32, _DP]),
    ("gpe_batch_chunk_node_steps", _U64, [_P, _U32]),
    ("gpe_batch_active", C.c_int32, [_P]),
    ("gpe_batch_resident_bytes", _U64, [_P]),
    ("gpe_batch_set_tile", _I, [_P, _U32]),
    ("gpe_batch_set_group_bytes", _I, [_P, _U64]),
    ("gpe_batch_group_bytes", _U64, [_P]),
    ("gpe_batch_set_launch_mode", _I, [_P, _U32]),
    ("gpe_server_create", _I, [_U32, _U64, _U32, C.POINTER(_P)]),
    ("gpe_server_destroy", None, [_P]),
    ("gpe_server_now", _D, [_P]),
    ("gpe_server_submit", _I, [_P, C.POINTER(GpeRequest), _D, C.POINTER(_U64)]),
    ("gpe_server_tick", _I, [_P, C.POINTER(_U32)]),
    ("gpe_server_poll", _I, [_P, C.POINTER(GpeCompletion), _U32, C.POINTER(_U32)]),
    ("gpe_server_fetch", _I, [_P, _U64, _DP, _DP]),
    ("gpe_server_get_stats", _I, [_P, C.POINTER(GpeServerStats)]),
    ("gpe_server_set_policy", _I, [_P, _U32, _U64]),
    ("gpe_server_oldest_queued", _D, [_P]),
    ("gpe_server_set_launch_mode", _I, [_P, _U32]),
    ("gpe_server_set_tile", _I, [_P, _U32]),
    ("gpe_server_cancel", _I, [_P, _U64]),
    ("gpe_checkpoint_save", _I, [_P, C.c_char_p]),
    ("gpe_checkpoint_load", _I, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("gpe_get_config", _I, [_P, _P]),
    ("gpe_server_set_timeout", _I, [_P, _D]),
    ("gpe_primal_create", _I, [C.POINTER(GpePrimalConfig), C.POINTER(_P)]),
    ("gpe_primal_destroy", None, [_P]),
    ("gpe_primal_set_state", _I, [_P, _DP, _DP]),
    ("gpe_primal_get_state", _I, [_P, _DP, _DP]),
    ("gpe_primal_set_mode", _I, [_P, _U32, _D]),
    ("gpe_primal_init_odd", _I, [_P, _D]),
    ("gpe_primal_step", _I, [_P, _U64]),
    ("gpe_primal_time_steps", _I, [_P, _U64, _DP, C.POINTER(_U64)]),
    ("gpe_primal_observables", _I, [_P, _I, _DP]),
    ("gpe_primal_synchronize", _I, [_P]),
    ("gpe_primal_tiles", _I, []),
    ("gpe_primal_coefficients", None, [_D, _D, _D, _DP]),
    ("gpe_primal_centered_origin", None, [_U32P, _D, _DP]),
    ("gpe_primal_positions", None, [_U32P, _DP, _D, _DP]),
    ("gpe_primal_trap_potential", None, [_U32P, _DP, _D, _D, _DP]),
    ("gpe_status_string", C.c_char_p, [_I]),
    ("gpe_last_error", C.c_char_p, []),
    ("gpe_bcc_positions", None, [_I32P]),
    ("gpe_bcc_stencil", None, [_I32P]),
    ("gpe_bcc_stencil_derived", _I, [_I32P]),
    ("gpe_vortex_state_load", _I, [C.c_char_p, C.POINTER(_P)]),
    ("gpe_vortex_state_free", None, [_P]),
    ("gpe_vortex_state_from_table", _I, [_U32, _D, _U32, _U32, _D, _D, _D, _D, _DP, C.POINTER(_P)]),
    ("gpe_vortex_state_table", _I, [_P, _DP]),
    ("gpe_vortex_state_params", None, [_P, _DP]),
    ("gpe_vortex_state_psi", None, [_P, _D, _D, _D, _DP]),
    ("gpe_vortex_state_phase", None, [_P, _D, _DP]),
    ("gpe_vortex_state_search_max", _D, [_P]),
    ("gpe_vortex_state_search_max_r", _D, [_P, _D]),
    ("gpe_vortex_state_search_max_z", _D, [_P, _D]),
    ("gpe_vortex_state_curvature", _D, [_P]),
    ("gpe_vortex_state_create", _I, [C.POINTER(_P)]),
    ("gpe_vortex_state_set_kappa", None, [_P, _U32]),
    ("gpe_vortex_state_set_g", None, [_P, _D]),
    ("gpe_vortex_state_set_threads", None, [_P, _U32]),
    ("gpe_vortex_state_set_dimensions", _I, [_P, _U32, _D, _D, _U32, _D]),
    ("gpe_vortex_state_set_range", _I, [_P, _D, _D, _D, _D, _D]),
    ("gpe_vortex_state_initialize", _I, [_P, _P]),
    ("gpe_vortex_state_iterate", _I, [_P, RZ_POTENTIAL, _D, _P, _U32, _D, RELAX_PROGRESS, _P,
                                      C.POINTER(GpeRelaxReport)]),
    ("gpe_vortex_state_function_error", _D, [_P, RZ_POTENTIAL, _D, _P]),
    ("gpe_vortex_state_function", _D, [_P, _D, _D]),
    ("gpe_vortex_state_search_min_r", _D, [_P, _D]),
    ("gpe_vortex_state_integrate_z", _D, [_P, _D]),
    ("gpe_vortex_state_save", _I, [_P, C.c_char_p]),
    ("gpe_vortex_state_reference_recipe", _I, [_U32, _D, _D, C.c_int32, _U32, _D, RELAX_PROGRESS, _P,
                                               C.POINTER(_P), _DP]),
    ("gpe_volume_map", _I, [_P, _D, _U32, C.POINTER(C.c_uint16)]),
    ("gpe_slice_picture", _I, [_P, _U32, _D, C.c_int32, C.POINTER(C.c_uint8)]),
    ("gpe_write_volume_map", _I, [C.c_char_p, C.POINTER(C.c_uint16), _U32P, _DP]),
    ("gpe_write_bmp", _I, [C.c_char_p, _U32, _U32, C.POINTER(C.c_uint8)]),
    ("gpe_integrate_in_time_output", _I, [_P, _D, _DP, _DP, _D, _U32, _D, _U32, C.POINTER(GpeOutputOptions),
                                          C.POINTER(GpeReferenceRun)]),
    ("gpe_reference_dims", None, [_D, _DP, _DP, _U32P, _DP]),
    ("gpe_reference_setup", _I, [_P, _D, _DP, _DP, _D, _D, _U32, C.POINTER(GpeReferenceRun), _DP, _DP, _DP]),
    ("gpe_integrate_in_time", _I, [_P, _D, _DP, _DP, _D, _U32, _D, _U32, C.POINTER(GpeReferenceRun), _DP, _DP]),
    ("gpe_centered_origin", None, [_U32P, _D, _DP]),
    ("gpe_node_positions", None, [_U32P, _DP, _D, _DP]),
    ("gpe_trap_potential", None, [_U32P, _DP, _D, _D, _DP]),
    ("gpe_field_norm", _D, [_U32P, _D, _DP]),
    ("gpe_recipe_gaussian", _I, [_U32P, _DP, _D, _D, _D, _U64, _DP]),
    ("gpe_recipe_thomas_fermi", _I, [_U32P, _DP, _D, _D, _D, _DP, _DP]),
    ("gpe_recipe_vortex_line", _I, [_U32P, _DP, _D, _D, _D, _DP]),
    ("gpe_recipe_vortex_ring", _I, [_U32P, _DP, _D, _DP, _DP, _D, _DP]),
    ("gpe_recipe_phase_shift", _I, [_U32P, _DP, _D, _D, _DP]),
]


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python vaikkari_amd/build.py` "
            "(hipcc --offload-arch=gfx950); there is no fallback implementation")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    msg = lib.gpe_last_error()
    return msg.decode() if msg else ""


def check(status: int, where: str) -> None:
    if status != OK:
        raise GpeError(status, where, f"{lib.gpe_status_string(status).decode()}: {last_error()}")


def dptr(a: np.ndarray):
    """Pointer to a C-contiguous float64 array (None passes NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_DP)


def u32(vals):
    return (C.c_uint32 * len(vals))(*[int(v) for v in vals])


def d3(vals):
    return (C.c_double * 3)(*[float(v) for v in vals])
