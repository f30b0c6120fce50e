"""ctypes binding of libpasta_hip.so (include/pasta_hip.h).

The library is the only compute path: if it is missing, importing fails loudly
(build it with `python pasta_amd/build.py`).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpasta_hip.so")

PASTA_OK = 0
PASTA_ERR_ARG = -1
PASTA_ERR_HIP = -2
PASTA_ERR_NOMEM = -3
PASTA_ERR_MODEL = -4
STATUS_TEXT = {
    0: "ok",
    1: "labels out of the grid bounds",
    2: "not enough neighbours",
    3: "vsini must be positive",
    4: "rotation kernel larger than FFTSIZE",
    5: "observed wavelength outside the synthetic range after the RV shift",
    6: "resample index past the end of the model",
    7: "particle is not within bounds",
    8: "continuum solve failed",
    9: "chi2 - target has the same sign at both ends of the search bracket",
    10: "search bound outside the grid",
}


class PastaError(RuntimeError):
    pass


class FitSettings(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int),
        ("max_iters", C.c_int),
        ("inertia_factor", C.c_double),
        ("cognitive_factor", C.c_double),
        ("social_factor", C.c_double),
        ("delta", C.c_double),
        ("vsini_range", C.c_double * 2),
        ("rv_range", C.c_double * 2),
        ("constraint_kind", C.c_int * 5),
        ("constraint_a", C.c_double * 5),
        ("constraint_b", C.c_double * 5),
    ]


class FitResult(C.Structure):
    _fields_ = [
        ("label", C.c_double * 5),
        ("chi2", C.c_double),
        ("iterations", C.c_int),
        ("status", C.c_int),
        ("time_s", C.c_double),
        ("n_params", C.c_int),
        ("continuum_params", C.c_float * 512),
    ]


class BinarySettings(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int),
        ("max_iters", C.c_int),
        ("inertia_factor", C.c_double),
        ("cognitive_factor", C.c_double),
        ("social_factor", C.c_double),
        ("delta", C.c_double),
        ("vsini_range", C.c_double * 2),
        ("rv_range", C.c_double * 2),
        ("light_ratio_range", C.c_double * 2),
        ("constraint_kind", C.c_int * 11),
        ("constraint_a", C.c_double * 11),
        ("constraint_b", C.c_double * 11),
    ]


class BinaryResult(C.Structure):
    _fields_ = [
        ("label1", C.c_double * 5),
        ("label2", C.c_double * 5),
        ("light_ratio", C.c_double),
        ("chi2", C.c_double),
        ("iterations", C.c_int),
        ("status", C.c_int),
        ("time_s", C.c_double),
        ("n_params", C.c_int),
        ("continuum_params", C.c_float * 512),
    ]


class TsResult(C.Structure):
    _fields_ = [
        ("label1", C.c_double * 4),
        ("label2", C.c_double * 4),
        ("light_ratio", C.c_double),
        ("chi2", C.c_double),
        ("iterations", C.c_int),
        ("status", C.c_int),
        ("time_s", C.c_double),
        ("n_params", C.c_int),
        ("n_spectra", C.c_int),
    ]


class RvResult(C.Structure):
    _fields_ = [
        ("rv1", C.c_double),
        ("rv2", C.c_double),
        ("chi2", C.c_double),
        ("iterations", C.c_int),
        ("status", C.c_int),
        ("time_s", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("max_stars", C.c_int),
        ("max_particles", C.c_int),
        ("iters_per_step", C.c_int),
        ("hbm_budget_bytes", C.c_int64),
        ("timing", C.c_int),
    ]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u16p = C.POINTER(C.c_uint16)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)

# name -> argtypes; every function returns int
SIGNATURES = {
    "pasta_version": [],
    "pasta_device_mem_info": [_u64p, _u64p],
    "pasta_grid_create": [_u16p, _fp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _PP],
    "pasta_grid_destroy": [_P],
    "pasta_grid_shape": [_P, _ip],
    "pasta_grid_axes": [_P, _dp, _dp, _dp, _ip, _ip, _ip],
    "pasta_grid_device_rows": [_P, C.POINTER(_u16p), _i64p],
    "pasta_grid_set_factors": [_P, _fp],
    "pasta_grid_set_fft_taps": [_P, C.c_int],
    "pasta_grid_local_weights": [_P, C.c_double, C.c_double, C.c_double, _ip, _fp],
    "pasta_grid_clamp_1d": [_P, C.c_double, C.c_double, C.c_double, C.c_int, _dp],
    "pasta_grid_is_teff_logg_between_bounds": [_P, C.c_double, C.c_double],
    "pasta_disp_create_none": [_P, _dp, C.c_int, _PP],
    "pasta_disp_create_fixed": [_P, _dp, C.c_int, C.c_double, _PP],
    "pasta_disp_create_variable": [_P, _dp, C.c_int, _fp, _PP],
    "pasta_disp_destroy": [_P],
    "pasta_disp_kernel": [_P, _fp, C.c_int64, _ip],
    "pasta_cont_create_chunk": [_dp, C.c_int, C.c_int, C.c_int, C.c_double, _PP],
    "pasta_cont_create_linear": [_fp, C.c_int, C.c_int, _PP],
    "pasta_cont_create_fixed": [_fp, C.c_int, C.c_int, _PP],
    "pasta_cont_create_constant": [C.c_int, _PP],
    "pasta_cont_destroy": [_P],
    "pasta_cont_n_params": [_P],
    "pasta_cont_chunks": [_P, _ip, C.c_int],
    "pasta_cont_build": [_P, _fp, C.c_int, _fp, C.c_int64],
    "pasta_star_create": [_P, _P, _P, _fp, _fp, C.c_int, _PP],
    "pasta_star_destroy": [_P],
    "pasta_star_eval_bytes": [_P, _i64p],
    "pasta_eval_chi2": [_P, _dp, C.c_int, _dp, _ip, _fp],
    "pasta_eval_chi2_noalloc": [_P, _dp, C.c_int, _dp, _ip, _fp],
    "pasta_produce_model": [_P, _P, _dp, C.c_int, _fp, _ip],
    "pasta_interpolate": [_P, _dp, C.c_int, _fp, _ip],
    "pasta_interpolate_and_convolve": [_P, _P, _dp, C.c_int, _fp, _ip],
    "pasta_vsini_kernel": [_P, C.c_double, _fp, C.c_int, _ip],
    "pasta_fit": [_P, C.POINTER(FitSettings), C.c_uint64, _dp, C.POINTER(FitResult), _dp],
    "pasta_fit_observed": [_P, C.POINTER(FitSettings), C.c_uint64, _dp, C.POINTER(FitResult), _dp, _dp],
    "pasta_fit_bulk": [_PP, C.c_int, C.POINTER(FitSettings), _u64p, C.POINTER(FitResult), _dp],
    "pasta_compute_uncertainty": [_P, C.c_double, _dp, _dp, _dp, _ip],
    "pasta_eval_chi2_multi": [_PP, C.c_int, _dp, _dp, _ip],
    "pasta_compute_uncertainty_bulk": [_PP, C.c_int, C.c_double, _dp, _dp, _dp, _ip],
    "pasta_server_create": [_P, C.POINTER(ServerConfig), _PP],
    "pasta_server_destroy": [_P],
    "pasta_server_submit": [_P, _P, C.POINTER(FitSettings), C.c_uint64, _u64p],
    "pasta_server_step": [_P, _ip, _ip],
    "pasta_server_poll": [_P, C.c_uint64, _ip, C.POINTER(FitResult)],
    "pasta_server_cancel": [_P, C.c_uint64],
    "pasta_server_stats": [_P, _i64p],
    "pasta_server_timing": [_P, C.c_int, _dp],
    "pasta_server_set_timing": [_P, C.c_int],
    "pasta_grid_read_rows": [_P, C.c_int, C.c_int, _u16p],
    "pasta_interpolate_linear": [_P, _dp, C.c_int, _fp, _ip],
    "pasta_binary_create": [_P, _P, _P, _P, _fp, _fp, C.c_int, _PP],
    "pasta_binary_destroy": [_P],
    "pasta_binary_model": [_P, _dp, C.c_int, _fp, _ip],
    "pasta_binary_eval_chi2": [_P, _dp, C.c_int, _dp, _ip, _fp],
    "pasta_binary_fit": [_P, C.POINTER(BinarySettings), C.c_uint64, _dp, C.POINTER(BinaryResult),
                         _dp],
    "pasta_binary_ts_create": [_P, _P, _P, _P, _fp, _fp, _dp, C.c_int, C.c_int, _PP],
    "pasta_binary_ts_eval_chi2": [_P, _dp, C.c_int, _dp, _ip],
    "pasta_binary_ts_fit": [_P, C.POINTER(BinarySettings), C.c_uint64, _dp, C.POINTER(TsResult),
                            _fp, _dp, _dp],
    "pasta_binary_rv_create": [_P, _P, _P, _fp, _fp, C.c_int, _fp, _fp, _fp, _fp, C.c_double, _PP],
    "pasta_binary_rv_destroy": [_P],
    "pasta_binary_rv_eval_chi2": [_P, _dp, C.c_int, _dp, _ip],
    "pasta_binary_rv_fit": [_P, C.POINTER(BinarySettings), C.c_uint64, _dp, C.POINTER(RvResult),
                            _dp],
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python pasta_amd/build.py` "
            "(there is no CPU fallback)"
        )
    lib = C.CDLL(LIB_PATH)
    lib.pasta_last_error.restype = C.c_char_p
    lib.pasta_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    return lib


lib = _load()


def last_error():
    return lib.pasta_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc < 0:
        raise PastaError(f"{what}: {last_error()}" if what else last_error())
    return rc


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
