"""ctypes wrapper of the C oracle (oracle/libpasta_oracle.so).

TEST INFRASTRUCTURE, NOT THE PRODUCT: imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, always as the checker
of the HIP path, never as the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpasta_oracle.so")

DISP_NONE, DISP_FIXED, DISP_VARIABLE = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()
_L = C.CDLL(LIB_PATH)

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u16p = C.POINTER(C.c_uint16)


class _Grid(C.Structure):
    _fields_ = [("n_teff", C.c_int), ("n_m", C.c_int), ("n_logg", C.c_int),
                ("teff", _dp), ("m", _dp), ("logg", _dp),
                ("logg_lo", _ip), ("logg_hi", _ip), ("cum", _ip), ("n_models", C.c_int)]


class _WlGrid(C.Structure):
    _fields_ = [("first", C.c_double), ("step", C.c_double), ("n", C.c_int)]


class _Model(C.Structure):
    _fields_ = [("g", C.c_void_p), ("spectra", _u16p), ("factors", _fp), ("sg", _WlGrid)]


class _Bounds(C.Structure):
    _fields_ = [("g", C.c_void_p), ("vsini_lo", C.c_double), ("vsini_hi", C.c_double),
                ("rv_lo", C.c_double), ("rv_hi", C.c_double), ("ckind", C.c_int * 5),
                ("ca", C.c_double * 5), ("cb", C.c_double * 5)]


class _PsoSettings(C.Structure):
    _fields_ = [("num_particles", C.c_int), ("max_iters", C.c_int), ("inertia", C.c_double),
                ("cognitive", C.c_double), ("social", C.c_double), ("delta", C.c_double)]


class _BinBounds(C.Structure):
    _fields_ = [("g", C.c_void_p), ("vsini_lo", C.c_double), ("vsini_hi", C.c_double),
                ("rv_lo", C.c_double), ("rv_hi", C.c_double), ("lr_lo", C.c_double),
                ("lr_hi", C.c_double), ("ckind", C.c_int * 11), ("ca", C.c_double * 11),
                ("cb", C.c_double * 11)]


COST_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, _dp, _dp)

SCALAR_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double))

_sig = {
    "or_grid_from_labels": (C.c_void_p, [_dp, C.c_int]),
    "or_grid_free": (None, [C.c_void_p]),
    "or_grid_model_index": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "or_grid_list_gridpoints": (C.c_int, [C.c_void_p, _dp]),
    "or_is_teff_logg_between_bounds": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "or_find_neighbors": (C.c_int, [_dp, C.c_int, C.c_double, C.c_int, C.c_int, _ip]),
    "or_local_weights": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, _ip, _fp, _dp,
                                   _dp, _dp, _dp]),
    "or_clamp_1d": (C.c_int, [C.c_void_p, _dp, C.c_int, _dp]),
    "or_interpolate_f32": (C.c_int, [C.c_void_p, _u16p, _fp, C.c_int, C.c_double, C.c_double,
                                     C.c_double, _fp]),
    "or_interpolate_f64": (C.c_int, [C.c_void_p, _u16p, _fp, C.c_int, C.c_double, C.c_double,
                                     C.c_double, _dp]),
    "or_rotation_kernel": (C.c_int, [C.c_double, C.c_double, _fp, C.c_int]),
    "or_rotation_kernel_f64": (C.c_int, [C.c_double, C.c_double, _dp, C.c_int]),
    "or_oa_convolve": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, C.c_int]),
    "or_oa_out_len": (C.c_int, [C.c_int, C.c_int]),
    "or_make_gaussian": (None, [_fp, C.c_int, C.c_float]),
    "or_fixed_gauss_kernel": (C.c_int, [C.c_double, C.c_double, _fp, C.c_int]),
    "or_direct_convolve_f64": (None, [_dp, C.c_int, _dp, C.c_int, _dp]),
    "or_disp_new": (C.c_void_p, [C.c_int, _dp, C.c_int, C.POINTER(_WlGrid), C.c_double, _fp]),
    "or_disp_free": (None, [C.c_void_p]),
    "or_disp_kernel_len": (C.c_int, [C.c_void_p]),
    "or_disp_kernels": (C.c_long, [C.c_void_p, _fp, C.c_long]),
    "or_cont_new_chunk": (C.c_void_p, [_dp, C.c_int, C.c_int, C.c_int, C.c_double]),
    "or_cont_new_linear": (C.c_void_p, [_fp, C.c_int, C.c_int]),
    "or_cont_new_fixed": (C.c_void_p, [_fp, C.c_int, C.c_int]),
    "or_cont_new_constant": (C.c_void_p, [C.c_int]),
    "or_cont_free": (None, [C.c_void_p]),
    "or_cont_n_params": (C.c_int, [C.c_void_p]),
    "or_cont_startstop": (C.c_int, [C.c_void_p, _ip]),
    "or_cont_design_row": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp]),
    "or_cont_fac": (C.c_int, [C.c_void_p, C.c_int, _fp, C.c_int]),
    "or_fit_continuum": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp, _dp]),
    "or_fit_continuum_f64": (C.c_int, [C.c_void_p, _fp, _fp, _dp, _dp, _dp]),
    "or_build_continuum": (C.c_int, [C.c_void_p, _fp, _fp]),
    "or_produce_model": (C.c_int, [C.POINTER(_Model), C.c_void_p, _dp, _fp]),
    "or_produce_model_f64": (C.c_int, [C.POINTER(_Model), C.c_void_p, _dp, _dp]),
    "or_produce_model_f64q": (C.c_int, [C.POINTER(_Model), C.c_void_p, _dp, _dp]),
    "or_interpolate_and_convolve": (C.c_int, [C.POINTER(_Model), C.c_void_p, _dp, _fp]),
    "or_cost": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _dp, _dp, _fp]),
    "or_cost_batch": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _dp, C.c_int,
                                C.c_int, _dp]),
    "or_cost_f64": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _dp, _dp]),
    "or_cost_batch_f64q": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _dp,
                                     C.c_int, C.c_int, _dp]),
    "or_cost_f64q": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _dp, _dp,
                               _dp]),
    "or_rng_u64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "or_rng_uniform": (C.c_double, [C.c_uint64, C.c_uint64]),
    "or_bounds_outer": (C.c_int, [C.POINTER(_Bounds), _dp, _dp]),
    "or_bounds_within": (C.c_int, [C.POINTER(_Bounds), _dp]),
    "or_bounds_clamp": (C.c_int, [C.POINTER(_Bounds), _dp, C.c_int, _dp]),
    "or_pso_init_positions": (C.c_int, [C.POINTER(_Bounds), C.c_int, C.c_uint64, _dp, _dp]),
    "or_pso_run": (C.c_int, [C.POINTER(_Bounds), C.POINTER(_PsoSettings), C.c_uint64, _dp, _dp,
                             COST_FN, C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_long)]),
    "or_fit": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, C.POINTER(_Bounds),
                         C.POINTER(_PsoSettings), C.c_uint64, _dp, C.c_int, _dp, _dp, _fp,
                         C.POINTER(C.c_long)]),
    "or_local_weights_linear": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, _ip,
                                          _fp, _dp]),
    "or_interpolate_linear_f32": (C.c_int, [C.c_void_p, _u16p, _fp, C.c_int, C.c_double,
                                            C.c_double, C.c_double, _fp]),
    "or_interpolate_linear_f64": (C.c_int, [C.c_void_p, _u16p, _fp, C.c_int, C.c_double,
                                            C.c_double, C.c_double, _dp]),
    "or_binary_model_norm": (C.c_int, [C.POINTER(_Model), C.POINTER(_Model), C.c_void_p, _dp, _dp,
                                       C.c_float, _fp]),
    "or_binary_cost": (C.c_int, [C.POINTER(_Model), C.POINTER(_Model), C.c_void_p, C.c_void_p,
                                 _fp, _fp, _dp, _dp, _fp]),
    "or_binbounds_outer": (C.c_int, [C.POINTER(_BinBounds), _dp, _dp]),
    "or_binbounds_within": (C.c_int, [C.POINTER(_BinBounds), _dp]),
    "or_binbounds_clamp": (C.c_int, [C.POINTER(_BinBounds), _dp, C.c_int, _dp]),
    "or_pso_run_binary": (C.c_int, [C.POINTER(_BinBounds), C.POINTER(_PsoSettings), C.c_uint64,
                                    _dp, COST_FN, C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_long)]),
    "or_pso_run_d": (C.c_int, [C.POINTER(_BinBounds), C.c_int, C.POINTER(_PsoSettings), C.c_uint64,
                               _dp, COST_FN, C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_long)]),
    "or_binary_rv_cost": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp, _fp, _fp,
                                    _fp, _fp, C.c_double, C.c_double, C.c_double, _dp]),
    "or_binary_ts_cost": (C.c_int, [C.POINTER(_Model), C.POINTER(_Model), C.c_void_p, C.c_void_p,
                                    _fp, _fp, _dp, C.c_int, _dp, _dp]),
    "or_brent_root": (C.c_int, [SCALAR_FN, C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _ip]),
    "or_compute_uncertainty": (C.c_int, [C.POINTER(_Model), C.c_void_p, C.c_void_p, _fp, _fp,
                                         C.c_double, _dp, _dp, _dp, _ip]),
}
for _n, (_r, _a) in _sig.items():
    _f = getattr(_L, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"oracle error {code} {what}")
        self.code = code


def _chk(rc, what=""):
    if rc < 0:
        raise OracleError(rc, what)
    return rc


class Grid:
    """Grid geometry + spectra (n_models x W uint16 in list_gridpoints order)."""

    def __init__(self, labels, spectra=None, factors=None, log10_first=None, log10_step=None):
        labels = _d(labels).reshape(-1, 3)
        self._h = _L.or_grid_from_labels(_p(labels, C.c_double), len(labels))
        if not self._h:
            raise OracleError(-7, "invalid grid labels")
        self._s = C.cast(self._h, C.POINTER(_Grid)).contents
        self.n_models = self._s.n_models
        self.spectra = None if spectra is None else np.ascontiguousarray(spectra, dtype=np.uint16)
        self.factors = None if factors is None else _f(factors)
        self.W = None if spectra is None else self.spectra.shape[1]
        self.sg = None
        if log10_first is not None:
            self.sg = _WlGrid(log10_first, log10_step, self.W or 0)

    def __del__(self):
        if getattr(self, "_h", None):
            _L.or_grid_free(self._h)
            self._h = None

    def axes(self):
        s = self._s
        return (np.array(s.teff[:s.n_teff]), np.array(s.m[:s.n_m]), np.array(s.logg[:s.n_logg]),
                np.array(s.logg_lo[:s.n_teff]), np.array(s.logg_hi[:s.n_teff]),
                np.array(s.cum[:s.n_teff]))

    def list_gridpoints(self):
        out = np.zeros((self.n_models, 3))
        _L.or_grid_list_gridpoints(self._h, _p(out, C.c_double))
        return out

    def local_weights(self, teff, m, logg):
        idx = np.zeros(64, np.int32)
        w32 = np.zeros(64, np.float32)
        w64 = np.zeros(64)
        ft = np.zeros(16); fl = np.zeros(16); fm = np.zeros(4)
        rc = _L.or_local_weights(self._h, teff, m, logg, _p(idx, C.c_int), _p(w32, C.c_float),
                                 _p(w64, C.c_double), _p(ft, C.c_double), _p(fl, C.c_double),
                                 _p(fm, C.c_double))
        _chk(rc, "local_weights")
        return idx, w32, w64, ft.reshape(4, 4), fl.reshape(4, 4), fm

    def local_weights_linear(self, teff, m, logg):
        idx = np.zeros(8, np.int32)
        w32 = np.zeros(8, np.float32)
        w64 = np.zeros(8)
        _chk(_L.or_local_weights_linear(self._h, teff, m, logg, _p(idx, C.c_int),
                                        _p(w32, C.c_float), _p(w64, C.c_double)),
             "local_weights_linear")
        return idx, w32, w64

    def interpolate_linear(self, teff, m, logg, f64=False):
        fac = None if self.factors is None else _p(self.factors, C.c_float)
        if f64:
            out = np.zeros(self.W)
            _chk(_L.or_interpolate_linear_f64(self._h, _p(self.spectra, C.c_uint16), fac, self.W,
                                              teff, m, logg, _p(out, C.c_double)),
                 "interpolate_linear")
        else:
            out = np.zeros(self.W, np.float32)
            _chk(_L.or_interpolate_linear_f32(self._h, _p(self.spectra, C.c_uint16), fac, self.W,
                                              teff, m, logg, _p(out, C.c_float)),
                 "interpolate_linear")
        return out

    def model(self):
        fac = None if self.factors is None else _p(self.factors, C.c_float)
        return _Model(self._h, _p(self.spectra, C.c_uint16), fac, self.sg)

    def interpolate(self, teff, m, logg, f64=False):
        fac = None if self.factors is None else _p(self.factors, C.c_float)
        if f64:
            out = np.zeros(self.W)
            _chk(_L.or_interpolate_f64(self._h, _p(self.spectra, C.c_uint16), fac, self.W, teff, m,
                                       logg, _p(out, C.c_double)), "interpolate")
        else:
            out = np.zeros(self.W, np.float32)
            _chk(_L.or_interpolate_f32(self._h, _p(self.spectra, C.c_uint16), fac, self.W, teff, m,
                                       logg, _p(out, C.c_float)), "interpolate")
        return out


class Dispersion:
    def __init__(self, grid, kind, wl, resolution=0.0, disp=None):
        wl = _d(wl)
        self.wl = wl
        self.n = len(wl)
        self.kind = kind
        dp = None if disp is None else _p(_f(disp), C.c_float)
        self._keep = None if disp is None else _f(disp)
        if self._keep is not None:
            dp = _p(self._keep, C.c_float)
        self._h = _L.or_disp_new(kind, _p(wl, C.c_double), len(wl), C.byref(grid.sg), resolution, dp)
        if not self._h:
            raise OracleError(-7, "dispersion")

    def __del__(self):
        if getattr(self, "_h", None):
            _L.or_disp_free(self._h)
            self._h = None

    @property
    def K(self):
        return _L.or_disp_kernel_len(self._h)

    def kernels(self, W):
        """fixed: the K taps; variable: the W x K per-pixel kernels."""
        K = self.K
        out = np.zeros(max(1, K * W), np.float32)
        n = _L.or_disp_kernels(self._h, _p(out, C.c_float), len(out))
        if n < 0:
            raise OracleError(n, "kernels")
        return out[:K] if self.kind == DISP_FIXED else out[:n].reshape(W, K)


class Continuum:
    def __init__(self, kind, n, wl=None, n_chunks=5, p_order=8, overlap=0.2, design=None,
                 fixed=None, ignore=0):
        self.kind = kind
        self.n = n
        if kind == "chunk":
            wl = _d(wl)
            self._h = _L.or_cont_new_chunk(_p(wl, C.c_double), n, n_chunks, p_order, overlap)
        elif kind == "linear":
            self._design = _f(design)
            self._h = _L.or_cont_new_linear(_p(self._design, C.c_float), n, self._design.shape[1])
        elif kind == "fixed":
            self._fixed = _f(fixed)
            self._h = _L.or_cont_new_fixed(_p(self._fixed, C.c_float), n, ignore)
        else:
            self._h = _L.or_cont_new_constant(n)
        if not self._h:
            raise OracleError(-7, "continuum")

    def __del__(self):
        if getattr(self, "_h", None):
            _L.or_cont_free(self._h)
            self._h = None

    @property
    def n_params(self):
        return _L.or_cont_n_params(self._h)

    def startstop(self):
        out = np.zeros(64, np.int32)
        k = _chk(_L.or_cont_startstop(self._h, _p(out, C.c_int)))
        return out[: 2 * k].reshape(k, 2)

    def design_row(self, chunk, row, p1):
        out = np.zeros(p1, np.float32)
        _chk(_L.or_cont_design_row(self._h, chunk, row, _p(out, C.c_float)))
        return out

    def fac(self, boundary, cap=10_000_000):
        out = np.zeros(cap, np.float32)
        k = _chk(_L.or_cont_fac(self._h, boundary, _p(out, C.c_float), cap))
        return out[:k]

    def fit(self, flux, var, synth):
        flux, var = _f(flux), _f(var)
        params = np.zeros(max(1, self.n_params), np.float32)
        chi2 = C.c_double()
        if np.asarray(synth).dtype == np.float64:
            p64 = np.zeros(max(1, self.n_params))
            s = _d(synth)
            _chk(_L.or_fit_continuum_f64(self._h, _p(flux, C.c_float), _p(var, C.c_float),
                                         _p(s, C.c_double), _p(p64, C.c_double), C.byref(chi2)))
            return p64, chi2.value
        s = _f(synth)
        _chk(_L.or_fit_continuum(self._h, _p(flux, C.c_float), _p(var, C.c_float),
                                 _p(s, C.c_float), _p(params, C.c_float), C.byref(chi2)))
        return params, chi2.value

    def build(self, params):
        params = _f(params)
        out = np.zeros(self.n, np.float32)
        k = _chk(_L.or_build_continuum(self._h, _p(params, C.c_float), _p(out, C.c_float)))
        return out[:k]


def rotation_kernel(vsini, dvelo, f64=False):
    if f64:
        out = np.zeros(4096)
        n = _chk(_L.or_rotation_kernel_f64(vsini, dvelo, _p(out, C.c_double), 4096))
    else:
        out = np.zeros(4096, np.float32)
        n = _chk(_L.or_rotation_kernel(vsini, dvelo, _p(out, C.c_float), 4096))
    return out[:n]


def oa_convolve(x, h):
    x, h = _f(x), _f(h)
    n = _L.or_oa_out_len(len(x), len(h))
    y = np.zeros(n, np.float32)
    _chk(_L.or_oa_convolve(_p(x, C.c_float), len(x), _p(h, C.c_float), len(h), _p(y, C.c_float), n))
    return y


def fixed_gauss_kernel(resolution, log_step):
    out = np.zeros(4096, np.float32)
    n = _chk(_L.or_fixed_gauss_kernel(resolution, log_step, _p(out, C.c_float), 4096))
    return out[:n]


def produce_model(grid, disp, labels, f64=False):
    """f64=False: the f32-faithful restatement; True: every step in FP64 with exact
    weights; "q": FP64 arithmetic over the values the reference stores in f32 (the 64
    interpolation weights, the rotation and instrumental taps, the resample weight)."""
    lab = _d(labels)
    M = grid.model()
    if f64:
        out = np.zeros(disp.n)
        fn = _L.or_produce_model_f64q if f64 == "q" else _L.or_produce_model_f64
        rc = fn(C.byref(M), disp._h, _p(lab, C.c_double), _p(out, C.c_double))
    else:
        out = np.zeros(disp.n, np.float32)
        rc = _L.or_produce_model(C.byref(M), disp._h, _p(lab, C.c_double), _p(out, C.c_float))
    _chk(rc, "produce_model")
    return out


def interpolate_and_convolve(grid, disp, labels4):
    lab = _d(labels4)
    M = grid.model()
    out = np.zeros(grid.W, np.float32)
    _chk(_L.or_interpolate_and_convolve(C.byref(M), disp._h, _p(lab, C.c_double),
                                        _p(out, C.c_float)), "interpolate_and_convolve")
    return out


def cost(grid, disp, cont, flux, var, labels, f64=False):
    """chi2 of one label vector; returns (chi2, params) or raises OracleError.  f64 as in
    produce_model ("q": FP64 over the reference's f32-stored values, params in f64)."""
    lab = _d(labels)
    flux, var = _f(flux), _f(var)
    M = grid.model()
    c = C.c_double()
    if f64 == "q":
        params = np.zeros(max(1, cont.n_params))
        _chk(_L.or_cost_f64q(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
                             _p(lab, C.c_double), C.byref(c), _p(params, C.c_double)), "cost")
        return c.value, params
    if f64:
        _chk(_L.or_cost_f64(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
                            _p(lab, C.c_double), C.byref(c)), "cost")
        return c.value, None
    params = np.zeros(max(1, cont.n_params), np.float32)
    _chk(_L.or_cost(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
                    _p(lab, C.c_double), C.byref(c), _p(params, C.c_float)), "cost")
    return c.value, params


def bounds(grid, vsini_range, rv_range, constraints=()):
    b = _Bounds()
    b.g = grid._h
    b.vsini_lo, b.vsini_hi = vsini_range
    b.rv_lo, b.rv_hi = rv_range
    for (d, kind, a, bb) in constraints:
        b.ckind[d] = kind
        b.ca[d] = a
        b.cb[d] = bb
    return b


def pso_settings(P, iters, inertia=0.7213475204, cognitive=1.1931471806, social=1.1931471806,
                 delta=1e-7):
    return _PsoSettings(P, iters, inertia, cognitive, social, delta)


def pso_run(b, s, seed, cost_fn, init_pos=None, init_vel=None):
    """PSO with a Python cost function f(positions[n,5]) -> costs[n]."""
    def _cb(ctx, n, pos, out):
        try:
            p = np.ctypeslib.as_array(pos, shape=(n, 5)).copy()
            c = np.asarray(cost_fn(p), dtype=np.float64)
            for i in range(n):
                out[i] = c[i]
            return 0
        except Exception:
            return -7
    cb = COST_FN(_cb)
    best = np.zeros(5)
    bc = C.c_double()
    trace = np.zeros((s.max_iters + 1, 6))
    ne = C.c_long()
    ip = None if init_pos is None else _p(_d(init_pos), C.c_double)
    iv = None if init_vel is None else _p(_d(init_vel), C.c_double)
    keep = (_d(init_pos) if init_pos is not None else None, _d(init_vel) if init_vel is not None else None)
    if keep[0] is not None:
        ip = _p(keep[0], C.c_double)
    if keep[1] is not None:
        iv = _p(keep[1], C.c_double)
    _chk(_L.or_pso_run(C.byref(b), C.byref(s), seed, ip, iv, cb, None, _p(best, C.c_double),
                       C.byref(bc), _p(trace, C.c_double), C.byref(ne)), "pso_run")
    return best, bc.value, trace, ne.value


def pso_init_positions(b, P, seed):
    pos = np.zeros((P, 5))
    vel = np.zeros((P, 5))
    _chk(_L.or_pso_init_positions(C.byref(b), P, seed, _p(pos, C.c_double), _p(vel, C.c_double)))
    return pos, vel


def fit(grid, disp, cont, flux, var, b, s, seed, init_pos=None, n_threads=1):
    flux, var = _f(flux), _f(var)
    M = grid.model()
    best = np.zeros(5)
    bc = C.c_double()
    params = np.zeros(max(1, cont.n_params), np.float32)
    ne = C.c_long()
    ipa = None if init_pos is None else _d(init_pos)
    ip = None if ipa is None else _p(ipa, C.c_double)
    _chk(_L.or_fit(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
                   C.byref(b), C.byref(s), seed, ip, n_threads, _p(best, C.c_double), C.byref(bc),
                   _p(params, C.c_float), C.byref(ne)), "fit")
    return best, bc.value, params, ne.value


def rng_uniform(key, ctr):
    return _L.or_rng_uniform(key, ctr)


def rng_u64(key, ctr):
    return _L.or_rng_u64(key, ctr)


def cost_batch(grid, disp, cont, flux, var, labels, n_threads=1, f64=False):
    """chi2 of many label vectors on n_threads threads (errors raise); f64="q" for the
    FP64 mode over the reference's f32-stored values."""
    lab = _d(labels).reshape(-1, 5)
    flux, var = _f(flux), _f(var)
    M = grid.model()
    out = np.zeros(len(lab))
    fn = _L.or_cost_batch_f64q if f64 == "q" else _L.or_cost_batch
    _chk(fn(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
            _p(lab, C.c_double), len(lab), n_threads, _p(out, C.c_double)), "cost_batch")
    return out


def brent_root(fn, lo, hi, tol):
    """argmin's BrentRoot restated: root of fn on [lo, hi]; returns (root, n_evals)."""
    def _cb(ctx, x, out):
        try:
            out[0] = float(fn(x))
            return 0
        except Exception:
            return -7
    cb = SCALAR_FN(_cb)
    r = C.c_double()
    n = C.c_int()
    _chk(_L.or_brent_root(cb, None, lo, hi, tol, C.byref(r), C.byref(n)), "brent_root")
    return r.value, n.value


def compute_uncertainty(grid, disp, cont, flux, var, spec_res, label, search_radius=None):
    """(out[5,2] lower/upper distances, status[5,2]) as compute_uncertainty (fitting.rs:482-568)."""
    flux, var = _f(flux), _f(var)
    M = grid.model()
    out = np.zeros(10)
    st = np.zeros(10, np.int32)
    sr = None if search_radius is None else _d(search_radius)
    _chk(_L.or_compute_uncertainty(C.byref(M), disp._h, cont._h, _p(flux, C.c_float),
                                   _p(var, C.c_float), spec_res, _p(_d(label), C.c_double),
                                   None if sr is None else _p(sr, C.c_double), _p(out, C.c_double),
                                   _p(st, C.c_int)), "compute_uncertainty")
    return out.reshape(5, 2), st.reshape(5, 2)


def binary_model_norm(grid, cgrid, disp, labels1, labels2, light_ratio):
    """produce_binary_model_norm (interpolate.rs:868-890): f32 model at the observed pixels."""
    M, CM = grid.model(), cgrid.model()
    out = np.zeros(disp.n, np.float32)
    _chk(_L.or_binary_model_norm(C.byref(M), C.byref(CM), disp._h, _p(_d(labels1), C.c_double),
                                 _p(_d(labels2), C.c_double), float(light_ratio),
                                 _p(out, C.c_float)), "binary_model_norm")
    return out


def binary_cost(grid, cgrid, disp, cont, flux, var, labels11):
    """BinaryCostFunction::cost (fitting.rs:581-604); returns (chi2, params)."""
    flux, var = _f(flux), _f(var)
    M, CM = grid.model(), cgrid.model()
    c = C.c_double()
    params = np.zeros(max(1, cont.n_params), np.float32)
    _chk(_L.or_binary_cost(C.byref(M), C.byref(CM), disp._h, cont._h, _p(flux, C.c_float),
                           _p(var, C.c_float), _p(_d(labels11), C.c_double), C.byref(c),
                           _p(params, C.c_float)), "binary_cost")
    return c.value, params


def binary_bounds(grid, vsini_range, rv_range, light_ratio_range=(0.0, 1.0), constraints=()):
    b = _BinBounds()
    b.g = grid._h
    b.vsini_lo, b.vsini_hi = vsini_range
    b.rv_lo, b.rv_hi = rv_range
    b.lr_lo, b.lr_hi = light_ratio_range
    for (d, kind, a, bb) in constraints:
        b.ckind[d] = kind
        b.ca[d] = a
        b.cb[d] = bb
    return b


def binary_bounds_outer(b):
    lo, hi = np.zeros(11), np.zeros(11)
    _L.or_binbounds_outer(C.byref(b), _p(lo, C.c_double), _p(hi, C.c_double))
    return lo, hi


def binary_bounds_within(b, p):
    return bool(_L.or_binbounds_within(C.byref(b), _p(_d(p), C.c_double)))


def binary_bounds_clamp(b, p, d):
    out = C.c_double()
    _chk(_L.or_binbounds_clamp(C.byref(b), _p(_d(p), C.c_double), d, C.byref(out)), "clamp")
    return out.value


def binary_rv_cost(grid, disp, cont, flux, var, model1, model2, cont1, cont2, light_ratio,
                   rv1, rv2):
    """RVCostFunction::cost (fitting.rs:736-781) on synthetic-grid components."""
    flux, var = _f(flux), _f(var)
    arrs = [_f(a) for a in (model1, model2, cont1, cont2)]
    M = grid.model()
    c = C.c_double()
    _chk(_L.or_binary_rv_cost(C.byref(M), disp._h, cont._h, _p(flux, C.c_float), _p(var, C.c_float),
                              *[_p(a, C.c_float) for a in arrs], float(light_ratio), float(rv1),
                              float(rv2), C.byref(c)), "binary_rv_cost")
    return c.value


def binary_ts_cost(grid, cgrid, disp, cont, fluxes, vars_, rvs, labels9):
    """BinaryTimeseriesKnownRVCostFunction::cost (fitting.rs:954-985)."""
    fl, vr = _f(np.stack(fluxes)), _f(np.stack(vars_))
    rv = _d(rvs).reshape(-1, 2)
    M, CM = grid.model(), cgrid.model()
    c = C.c_double()
    _chk(_L.or_binary_ts_cost(C.byref(M), C.byref(CM), disp._h, cont._h, _p(fl, C.c_float),
                              _p(vr, C.c_float), _p(rv, C.c_double), len(rv),
                              _p(_d(labels9), C.c_double), C.byref(c)), "binary_ts_cost")
    return c.value


def pso_run_d(b, D, s, seed, cost_fn, init_pos=None):
    """The binary PSO over D = 11, 9 or 2 parameters with a Python cost f(pos[n,D])."""
    def _cb(ctx, n, pos, out):
        try:
            p = np.ctypeslib.as_array(pos, shape=(n, D)).copy()
            c = np.asarray(cost_fn(p), dtype=np.float64)
            for i in range(n):
                out[i] = c[i]
            return 0
        except Exception:
            return -7
    cb = COST_FN(_cb)
    best = np.zeros(D)
    bc = C.c_double()
    trace = np.zeros((s.max_iters + 1, D + 1))
    ne = C.c_long()
    ipa = None if init_pos is None else _d(init_pos)
    _chk(_L.or_pso_run_d(C.byref(b), D, C.byref(s), seed,
                         None if ipa is None else _p(ipa, C.c_double), cb, None,
                         _p(best, C.c_double), C.byref(bc), _p(trace, C.c_double), C.byref(ne)),
         "pso_run_d")
    return best, bc.value, trace, ne.value


def pso_run_binary(b, s, seed, cost_fn, init_pos=None):
    """The 11-parameter PSO with a Python cost f(positions[n,11]) -> costs[n]."""
    def _cb(ctx, n, pos, out):
        try:
            p = np.ctypeslib.as_array(pos, shape=(n, 11)).copy()
            c = np.asarray(cost_fn(p), dtype=np.float64)
            for i in range(n):
                out[i] = c[i]
            return 0
        except Exception:
            return -7
    cb = COST_FN(_cb)
    best = np.zeros(11)
    bc = C.c_double()
    trace = np.zeros((s.max_iters + 1, 12))
    ne = C.c_long()
    ipa = None if init_pos is None else _d(init_pos)
    _chk(_L.or_pso_run_binary(C.byref(b), C.byref(s), seed,
                              None if ipa is None else _p(ipa, C.c_double), cb, None,
                              _p(best, C.c_double), C.byref(bc), _p(trace, C.c_double),
                              C.byref(ne)), "pso_run_binary")
    return best, bc.value, trace, ne.value
