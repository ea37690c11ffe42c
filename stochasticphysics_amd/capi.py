"""ctypes binding of the C ABI declared in include/stophys_gpu.h.

The library ``libstophys_gpu.so`` is built in-tree by ``stochasticphysics_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback behind this API.
"""
import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libstophys_gpu.so"

# enum values (identical to the reference's selector enums)
PROBLEM = dict(Neel=1, NeelSpherical=2, BrownAndNeel=3, BrownAndNeelEulerSpherical=4)
SOLVER = dict(EulerMaruyama=1, EulerMaruyama_Normalized=2, Implicit_Midpoint=3, Millstein=6, Heun_Strong=7,
              Heun_NotConsistent=8)
FIELD = dict(none=1, constant=2, sinusoidal=3, lissajous=4, triangular=5, rectangular=6, sinc=7)
ANISO = dict(uniaxial=1, cubic=2, uniaxialcubic=4)
PREC = dict(fp64=0, mixed=1)
NOISE = dict(philox=0, replay=1)
FLAG_BN_SIMPLE, FLAG_COORD_TRANSFORM, FLAG_KEEP_FINAL, FLAG_NO_FALLBACK, FLAG_SKIP_CHECK = 1, 2, 4, 8, 16
FLAG_RENORMALIZE = 32
FLAG_EULER_TRANSFORM = 64
DIM_Y = {1: 3, 2: 2, 3: 6, 4: 5}
DIM_W = {1: 3, 2: 3, 3: 6, 4: 6}
DIM_OUT = {1: 3, 2: 3, 3: 6, 4: 9}
NPARAM = 23
# rows of the parameter block (sp_types.h ParamIndex)
P_A, P_B, P_SIG, P_D, P_MU, P_K1, P_PT, P_C, P_CF, P_DD, P_AX, P_KU, P_U = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 19, 20

SP_OK, SP_ERR_NOT_CONVERGED, SP_ERR_TIMEOUT, SP_ERR_NOT_READY = 0, 5, 7, 9


class SpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Request(C.Structure):
    _fields_ = [("problem", C.c_int32), ("solver", C.c_int32), ("field", C.c_int32), ("anisotropy", C.c_int32),
                ("precision", C.c_int32), ("noise", C.c_int32), ("flags", C.c_uint32), ("newton_max_iter", C.c_uint32),
                ("n_particles", C.c_uint64), ("n_steps", C.c_uint64), ("oversampling", C.c_uint64),
                ("seed", C.c_uint64), ("dt", C.c_double), ("newton_tol", C.c_double), ("min_angle", C.c_double),
                ("check_tol", C.c_double), ("field_offset", C.c_double * 3), ("field_amplitude", C.c_double * 3),
                ("field_freq", C.c_double * 3), ("field_phase", C.c_double * 3),
                ("replay_normals", C.POINTER(C.c_double)), ("replay_count", C.c_uint64),
                ("field_period", C.c_double), ("field_time_offset", C.c_double), ("field_tau", C.c_double),
                ("field_alternating", C.c_uint32), ("save_interval", C.c_uint32), ("brown_min_angle", C.c_double)]


class Particles(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in (
        "magnetic_radius", "saturation_magnetisation", "damping", "gyromagnetic_ratio", "temperature",
        "anisotropy_constant", "hydrodynamic_radius", "viscosity", "axes", "initial_state",
        "anisotropy_constant_cubic", "uniaxial_rotation")]


class Result(C.Structure):
    _fields_ = [("time", C.POINTER(C.c_double)), ("field", C.POINTER(C.c_double)), ("mean", C.POINTER(C.c_double)),
                ("weight", C.POINTER(C.c_double)), ("final_state", C.POINTER(C.c_double)),
                ("trajectories", C.POINTER(C.c_double)), ("n_saved", C.c_uint64),
                ("n_out", C.c_uint64), ("dim_out", C.c_uint32), ("dim_y", C.c_uint32),
                ("device_ms", C.c_double), ("latency_ms", C.c_double), ("particle_steps", C.c_double),
                ("check_error", C.c_double), ("check_noise", C.c_double), ("used_fallback", C.c_uint32),
                ("precision_used", C.c_uint32),
                ("newton_failures", C.c_uint64), ("particle_means", C.c_double * 8),
                ("check_bias", C.c_double), ("check_bias_se", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("chunk_steps", C.c_uint32),
                ("use_graph", C.c_uint32), ("persistent_waves_per_simd", C.c_uint32)]


EXPORTS = [
    "sp_last_error", "sp_abi_version", "sp_parse_enum", "sp_enum_name", "sp_request_shape",
    "sp_compute_particle_params", "sp_ctx_create", "sp_ctx_destroy", "sp_ctx_set_tuning", "sp_ctx_get_tuning",
    "sp_ctx_set_memory_budget", "sp_ctx_last_run_stats", "sp_run", "sp_run_batch", "sp_submit", "sp_wait",
    "sp_result_get", "sp_cancel", "sp_server_stats", "sp_autotune", "sp_dev_philox", "sp_dev_normals", "sp_dev_box_muller", "sp_host_box_muller",
    "sp_host_normals", "sp_prepare", "sp_execute", "sp_release", "sp_settings_load", "sp_settings_request",
    "sp_settings_savefile", "sp_settings_write_result", "sp_settings_free", "sp_settings_single_savefile",
    "sp_settings_write_single_results", "sp_settings_prepare", "sp_settings_submit", "sp_prepared_inputs",
    "sp_ctx_set_reduction_limit", "sp_ctx_inject_failures", "sp_execute_noalloc", "sp_progress",
    "sp_ctx_set_iteration_budget",
]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m stochasticphysics_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        P, d = C.POINTER, C.POINTER(C.c_double)
        L.sp_last_error.restype = C.c_char_p
        L.sp_enum_name.restype = C.c_char_p
        L.sp_enum_name.argtypes = [C.c_char_p, C.c_int32]
        L.sp_parse_enum.argtypes = [C.c_char_p, C.c_char_p, P(C.c_int32)]
        L.sp_request_shape.argtypes = [P(Request), P(C.c_uint64), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.sp_compute_particle_params.argtypes = [P(Request), P(Particles), d, P(C.c_uint32)]
        L.sp_ctx_create.argtypes = [C.c_int, P(C.c_void_p)]
        L.sp_ctx_destroy.argtypes = [C.c_void_p]
        L.sp_ctx_destroy.restype = None
        L.sp_ctx_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(Tuning)]
        L.sp_ctx_get_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(Tuning)]
        L.sp_ctx_set_memory_budget.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32]
        L.sp_ctx_set_reduction_limit.argtypes = [C.c_void_p, C.c_uint64]
        L.sp_ctx_set_iteration_budget.argtypes = [C.c_void_p, C.c_double]
        L.sp_ctx_inject_failures.argtypes = [C.c_void_p, C.c_uint32]
        L.sp_execute_noalloc.argtypes = [C.c_int, C.c_void_p, P(Request), C.c_void_p, C.c_void_p, C.c_void_p]
        L.sp_ctx_last_run_stats.argtypes = [C.c_void_p, d]
        L.sp_run.argtypes = [C.c_void_p, P(Request), P(Particles), P(Result)]
        L.sp_run_batch.argtypes = [C.c_void_p, C.c_uint32, P(Request), P(Particles), P(Result)]
        L.sp_submit.argtypes = [C.c_void_p, P(Request), P(Particles), P(C.c_uint64)]
        L.sp_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
        L.sp_result_get.argtypes = [C.c_void_p, C.c_uint64, P(Result)]
        L.sp_cancel.argtypes = [C.c_void_p, C.c_uint64]
        L.sp_progress.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]
        L.sp_server_stats.argtypes = [C.c_void_p, d]
        L.sp_autotune.argtypes = [C.c_void_p, P(Request), P(Particles), P(Tuning), d]
        L.sp_dev_philox.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_uint32), C.c_uint64, P(C.c_uint32)]
        L.sp_dev_normals.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, d]
        L.sp_prepare.argtypes = [C.c_void_p, P(Request), P(Particles), P(C.c_void_p)]
        L.sp_execute.argtypes = [C.c_void_p, P(Result)]
        L.sp_release.argtypes = [C.c_void_p]
        L.sp_release.restype = None
        L.sp_settings_load.argtypes = [C.c_char_p, P(C.c_void_p)]
        L.sp_settings_request.argtypes = [C.c_void_p, P(Request), P(Particles)]
        L.sp_settings_savefile.argtypes = [C.c_void_p]
        L.sp_settings_savefile.restype = C.c_char_p
        L.sp_settings_write_result.argtypes = [C.c_void_p, P(Result), C.c_char_p]
        L.sp_settings_single_savefile.argtypes = [C.c_void_p]
        L.sp_settings_single_savefile.restype = C.c_char_p
        L.sp_settings_write_single_results.argtypes = [C.c_void_p, P(Result), C.c_char_p]
        L.sp_settings_prepare.argtypes = [C.c_void_p, C.c_void_p, P(C.c_void_p)]
        L.sp_settings_submit.argtypes = [C.c_void_p, C.c_void_p, P(C.c_uint64)]
        L.sp_prepared_inputs.argtypes = [C.c_void_p, d, d]
        L.sp_settings_free.argtypes = [C.c_void_p]
        L.sp_settings_free.restype = None
        L.sp_host_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, d]
        L.sp_dev_box_muller.argtypes = [C.c_void_p, P(C.c_uint32), C.c_uint64, C.c_int, d]
        L.sp_host_box_muller.argtypes = [P(C.c_uint32), C.c_uint64, d]
        _LIB = L
    return _LIB


def check(rc):
    if rc != SP_OK:
        raise SpError(rc, lib().sp_last_error().decode())
    return rc


def dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
