"""Python mirror of the reference's PyO3 module `pasta` (src/lib.rs) over libpasta_hip.so.

Same names, argument meaning and error behaviour as the reference's Python API
for the model-evaluation path:

    WlGrid, InMemInterpolator, PSOSettings, Label, OptimizationResult,
    NoConvolutionDispersion, FixedResolutionDispersion, VariableResolutionDispersion,
    ChunkContinuumFitter, LinearModelContinuumFitter, FixedContinuum, ConstantContinuum,
    FixConstraint, RangeConstraint, get_vsini_kernel,
    InMemInterpolator.{produce_model, produce_model_bulk, interpolate,
        interpolate_and_convolve, get_fitter, calculate_interpolation_coefficients,
        get_neighbor_indices, teffs, ms, loggs, logg_limits, cumulative_grid_size,
        list_gridpoints, is_teff_logg_between_bounds, clamp_1d}
    InMemSingleFitter.{fit, fit_bulk, chi2}
    InMemInterpolator.{interpolate_linear, produce_binary_model_norm, get_binary_fitter,
        get_binary_rv_fitter, get_timeseries_known_rv_fitter}
    BinaryFitter, BinaryRVFitter, BinaryTimeseriesKnownRVFitter (fit, chi2),
    BinaryOptimizationResult, RVOptimizationResult, RvLessLabel,
    BinaryTimeseriesOptimizationResult

Every computation runs in the HIP library; there is no CPU fallback.  Errors
raise ValueError like PyO3's anyhow conversion.
"""
import ctypes as C
import io
import json
import os

import numpy as np

from . import _lib
from ._lib import check, f32, f64, lib, ptr

__all__ = [
    "WlGrid", "InMemInterpolator", "PSOSettings", "Label", "OptimizationResult",
    "NoConvolutionDispersion", "FixedResolutionDispersion", "VariableResolutionDispersion",
    "ChunkContinuumFitter", "LinearModelContinuumFitter", "FixedContinuum", "ConstantContinuum",
    "FixConstraint", "RangeConstraint", "get_vsini_kernel", "InMemSingleFitter", "FitServer",
    "LabelUncertainties", "BinaryFitter", "BinaryOptimizationResult", "BinaryRVFitter",
    "RVOptimizationResult", "BinaryTimeseriesKnownRVFitter", "RvLessLabel",
    "BinaryTimeseriesOptimizationResult",
]


def _raise(rc, what):
    if rc < 0:
        raise ValueError(f"{what}: {_lib.last_error()}")
    return rc


def _batch_rc(rc, what):
    """Return code of a batched call whose per-item failures are in its status array:
    PASTA_ERR_MODEL passes (the statuses say which items failed), a bad argument raises
    ValueError, and a failure of the call itself (HIP error, allocation) raises
    PastaError, since no status was written."""
    if rc == 0 or rc == _lib.PASTA_ERR_MODEL:
        return rc
    if rc == _lib.PASTA_ERR_ARG:
        raise ValueError(f"{what}: {_lib.last_error()}")
    raise _lib.PastaError(f"{what} failed ({rc}): {_lib.last_error()}")


class WlGrid:
    """WlGrid(min, step, len, log=False) (src/lib.rs:630-653).  Only log-spaced
    synthetic grids are supported by the HIP path (every grid this path is built for is
    log-uniform); Linspace grids raise on use."""

    def __init__(self, min, step, len, log=None):
        self.min, self.step, self.len, self.log = float(min), float(step), int(len), bool(log)

    def get_array(self):
        i = np.arange(self.len)
        return 10.0 ** (self.min + self.step * i) if self.log else self.min + self.step * i


class PSOSettings:
    """PSOSettings(num_particles, max_iters, inertia, cognitive, social, delta) (src/lib.rs:42-72)."""

    def __init__(self, num_particles, max_iters, inertia_factor=None, cognitive_factor=None,
                 social_factor=None, delta=None):
        self.num_particles = int(num_particles)
        self.max_iters = int(max_iters)
        self.inertia_factor = 0.7213475204 if inertia_factor is None else float(inertia_factor)
        self.cognitive_factor = 1.1931471806 if cognitive_factor is None else float(cognitive_factor)
        self.social_factor = 1.1931471806 if social_factor is None else float(social_factor)
        self.delta = 1e-7 if delta is None else float(delta)


class Label:
    """Fitted labels (src/lib.rs:81-166)."""

    def __init__(self, teff, m, logg, vsini, rv):
        self.teff, self.m, self.logg, self.vsini, self.rv = map(float, (teff, m, logg, vsini, rv))

    def as_list(self):
        return [self.teff, self.m, self.logg, self.vsini, self.rv]

    def to_dict(self):
        return dict(teff=self.teff, m=self.m, logg=self.logg, vsini=self.vsini, rv=self.rv)

    @staticmethod
    def from_dict(d):
        return Label(d["teff"], d["m"], d["logg"], d["vsini"], d["rv"])


class OptimizationResult:
    """Output of a fit (src/lib.rs:212-294)."""

    def __init__(self, label, continuum_params, chi2, iterations, time):
        self.label, self.continuum_params = label, list(continuum_params)
        self.chi2, self.iterations, self.time = float(chi2), int(iterations), float(time)

    def to_dict(self):
        return dict(label=self.label.to_dict(), continuum_params=self.continuum_params,
                    chi2=self.chi2, iterations=self.iterations, time=self.time)

    def to_json(self):
        return json.dumps(self.to_dict())

    @staticmethod
    def from_dict(d):
        return OptimizationResult(Label.from_dict(d["label"]), d["continuum_params"], d["chi2"],
                                  d["iterations"], d["time"])

    @staticmethod
    def _from_c(r):
        lab = Label(*r.label[:5])
        return OptimizationResult(lab, r.continuum_params[: r.n_params], r.chi2, r.iterations,
                                  r.time_s)


class BinaryOptimizationResult:
    """Output of a binary fit (src/lib.rs:296-360)."""

    def __init__(self, label1, label2, light_ratio, continuum_params, chi2, iterations, time):
        self.label1, self.label2 = label1, label2
        self.light_ratio, self.continuum_params = float(light_ratio), list(continuum_params)
        self.chi2, self.iterations, self.time = float(chi2), int(iterations), float(time)

    def to_dict(self):
        return dict(label1=self.label1.to_dict(), label2=self.label2.to_dict(),
                    light_ratio=self.light_ratio, continuum_params=self.continuum_params,
                    chi2=self.chi2, iterations=self.iterations, time=self.time)

    def to_json(self):
        return json.dumps(self.to_dict())

    @staticmethod
    def from_dict(d):
        return BinaryOptimizationResult(Label.from_dict(d["label1"]), Label.from_dict(d["label2"]),
                                        d["light_ratio"], d["continuum_params"], d["chi2"],
                                        d["iterations"], d["time"])

    @staticmethod
    def _from_c(r):
        return BinaryOptimizationResult(Label(*r.label1[:5]), Label(*r.label2[:5]), r.light_ratio,
                                        r.continuum_params[: r.n_params], r.chi2, r.iterations,
                                        r.time_s)


class RVOptimizationResult:
    """Output of a binary RV fit (src/lib.rs:420-447)."""

    def __init__(self, rv1, rv2, chi2, iterations, time):
        self.rv1, self.rv2, self.chi2 = float(rv1), float(rv2), float(chi2)
        self.iterations, self.time = int(iterations), float(time)

    def to_dict(self):
        return dict(rv1=self.rv1, rv2=self.rv2, chi2=self.chi2, iterations=self.iterations,
                    time=self.time)

    def to_json(self):
        return json.dumps(self.to_dict())


class RvLessLabel:
    """(Teff, [M/H], logg, vsini) of one star (src/lib.rs:449-520)."""

    def __init__(self, teff, m, logg, vsini):
        self.teff, self.m, self.logg, self.vsini = map(float, (teff, m, logg, vsini))

    def as_list(self):
        return [self.teff, self.m, self.logg, self.vsini]

    def to_dict(self):
        return dict(teff=self.teff, m=self.m, logg=self.logg, vsini=self.vsini)

    @staticmethod
    def from_dict(d):
        return RvLessLabel(d["teff"], d["m"], d["logg"], d["vsini"])


class BinaryTimeseriesOptimizationResult:
    """Output of a time-series binary fit (src/lib.rs:522-600)."""

    def __init__(self, label1, label2, light_ratio, continuum_params, chis, iters, time):
        self.label1, self.label2, self.light_ratio = label1, label2, float(light_ratio)
        self.continuum_params = [list(map(float, p)) for p in continuum_params]
        self.chis, self.iters, self.time = list(map(float, chis)), int(iters), float(time)

    def to_dict(self):
        return dict(label1=self.label1.to_dict(), label2=self.label2.to_dict(),
                    light_ratio=self.light_ratio, continuum_params=self.continuum_params,
                    chis=self.chis, iters=self.iters, time=self.time)

    def to_json(self):
        return json.dumps(self.to_dict())

    @staticmethod
    def from_dict(d):
        return BinaryTimeseriesOptimizationResult(
            RvLessLabel.from_dict(d["label1"]), RvLessLabel.from_dict(d["label2"]),
            d["light_ratio"], d["continuum_params"], d["chis"], d["iters"], d["time"])


class LabelUncertainties:
    """Lower / upper uncertainty per label (src/lib.rs:170-210): (lower, upper), each a
    float or None where the search found no crossing."""
    NAMES = ("teff", "m", "logg", "vsini", "rv")

    def __init__(self, out, status):
        out = np.asarray(out, np.float64).reshape(5, 2)
        status = np.asarray(status).reshape(5, 2)
        for i, name in enumerate(self.NAMES):
            setattr(self, name, tuple(None if status[i, k] else float(out[i, k]) for k in range(2)))
        self.status = status

    def to_dict(self):
        return {n: getattr(self, n) for n in self.NAMES}

    def to_json(self):
        return json.dumps(self.to_dict())


class _Constraint:
    def __init__(self, parameter, kind, a, b):
        self.parameter, self.kind, self.a, self.b = int(parameter), kind, float(a), float(b)


def FixConstraint(parameter, value):
    """Fix one parameter during the fit (src/lib.rs:817-823)."""
    return _Constraint(parameter, 1, value, value)


def RangeConstraint(parameter, lower, upper):
    """Bound one parameter during the fit (src/lib.rs:829-835)."""
    return _Constraint(parameter, 2, lower, upper)


# ---------------------------------------------------------------- dispersion
class PyWavelengthDispersion:
    """Deferred dispersion description, bound to a grid's synthetic sampling on use
    (the reference binds it to `synth_wl` at construction, src/lib.rs:762-797)."""

    def __init__(self, kind, wl, synth_wl, resolution=None, disp=None):
        self.kind, self.wl, self.synth_wl = kind, f64(wl), synth_wl
        self.resolution, self.disp = resolution, None if disp is None else f32(disp)
        self._handles = {}

    def wavelength(self):
        return self.wl.tolist()

    def handle(self, grid):
        key = id(grid)
        if key not in self._handles:
            if not self.synth_wl.log:
                raise ValueError("only log-spaced synthetic grids are supported")
            h = C.c_void_p()
            n = len(self.wl)
            if self.kind == "none":
                rc = lib.pasta_disp_create_none(grid._h, ptr(self.wl, C.c_double), n, C.byref(h))
            elif self.kind == "fixed":
                rc = lib.pasta_disp_create_fixed(grid._h, ptr(self.wl, C.c_double), n,
                                                 float(self.resolution), C.byref(h))
            else:
                rc = lib.pasta_disp_create_variable(grid._h, ptr(self.wl, C.c_double), n,
                                                    ptr(self.disp, C.c_float), C.byref(h))
            _raise(rc, "dispersion")
            self._handles[key] = (grid, h)
        return self._handles[key][1]

    def __del__(self):
        for _, h in getattr(self, "_handles", {}).values():
            lib.pasta_disp_destroy(h)


def NoConvolutionDispersion(wl, wl_grid):
    return PyWavelengthDispersion("none", wl, wl_grid)


def FixedResolutionDispersion(wl, resolution, synth_wl):
    if resolution <= 0:
        raise ValueError("Resolution must be positive")
    return PyWavelengthDispersion("fixed", wl, synth_wl, resolution=resolution)


def VariableResolutionDispersion(wl, disp, synth_wl):
    return PyWavelengthDispersion("variable", wl, synth_wl, disp=disp)


# ----------------------------------------------------------------- continuum
class PyContinuumFitter:
    def __init__(self, kind, **kw):
        self.kind = kind
        h = C.c_void_p()
        if kind == "chunk":
            wl = f64(kw["wl"])
            self.n = len(wl)
            rc = lib.pasta_cont_create_chunk(ptr(wl, C.c_double), len(wl), kw["n_chunks"],
                                             kw["p_order"], kw["overlap"], C.byref(h))
        elif kind == "linear":
            dm = f32(kw["design_matrix"])
            self.n = dm.shape[0]
            rc = lib.pasta_cont_create_linear(ptr(dm, C.c_float), dm.shape[0], dm.shape[1],
                                              C.byref(h))
        elif kind == "fixed":
            cont = f32(kw["continuum"])
            self.n = len(cont)
            rc = lib.pasta_cont_create_fixed(ptr(cont, C.c_float), len(cont), kw["ignore"],
                                             C.byref(h))
        else:
            self.n = None
            rc = 0
            h = None
        _raise(rc, "continuum")
        self._h = h
        self._const = {}

    def handle(self, n):
        if self.kind != "constant":
            if n != self.n:
                raise ValueError("Length of observed and synthetic spectrum are different")
            return self._h
        if n not in self._const:
            h = C.c_void_p()
            _raise(lib.pasta_cont_create_constant(n, C.byref(h)), "continuum")
            self._const[n] = h
        return self._const[n]

    def n_params(self, n):
        return lib.pasta_cont_n_params(self.handle(n))

    def build_continuum(self, params):
        params = f32(params)
        n = self.n or 1
        out = np.zeros(max(n, 1), np.float32)
        k = _raise(lib.pasta_cont_build(self.handle(n), ptr(params, C.c_float), len(params),
                                        ptr(out, C.c_float), len(out)), "build_continuum")
        return out[:k]

    def __del__(self):
        if getattr(self, "_h", None):
            lib.pasta_cont_destroy(self._h)
        for h in getattr(self, "_const", {}).values():
            lib.pasta_cont_destroy(h)


def ChunkContinuumFitter(wl, n_chunks=None, p_order=None, overlap=None):
    """Chunked polynomials with sin^2 blending (src/lib.rs:944-960)."""
    return PyContinuumFitter("chunk", wl=wl, n_chunks=5 if n_chunks is None else n_chunks,
                             p_order=8 if p_order is None else p_order,
                             overlap=0.2 if overlap is None else overlap)


def LinearModelContinuumFitter(design_matrix):
    return PyContinuumFitter("linear", design_matrix=design_matrix)


def FixedContinuum(continuum, ignore_first_and_last=None):
    return PyContinuumFitter("fixed", continuum=continuum,
                             ignore=0 if ignore_first_and_last is None else ignore_first_and_last)


def ConstantContinuum():
    return PyContinuumFitter("constant")


# ---------------------------------------------------------------- interpolator
def label_from_filename(filename):
    """(Teff, [M/H], logg) of a grid file name such as lp0020_08000_0430.npy
    (label_from_filename, src/model_fetchers.rs:23-38): three '_'-separated fields of 6, 5
    and 4 characters before the first '.'; [M/H] is negative for an 'lm' prefix."""
    parts = filename.split(".")[0].split("_")
    if len(parts) != 3 or len(parts[0]) != 6 or len(parts[1]) != 5 or len(parts[2]) != 4:
        raise ValueError(f"Invalid filename: {filename}")
    sign = -1.0 if parts[0].startswith("lm") else 1.0
    try:
        return (float(int(parts[1])), sign * float(int(parts[0][2:])) / 100.0,
                float(int(parts[2])) / 100.0)
    except ValueError:
        raise ValueError(f"Invalid filename: {filename}") from None


def _npy_u16(buf, what):
    a = np.load(io.BytesIO(buf) if isinstance(buf, (bytes, bytearray)) else buf, allow_pickle=False)
    if a.dtype != np.uint16:
        raise ValueError(f"{what}: expected uint16")
    return a.ravel()


def read_spectrum(a, includes_factor):
    """read_spectrum (src/model_fetchers.rs:48-67): with includes_factor the first two
    uint16 values are the little-endian f32 factor (max), returned divided by 65535."""
    if not includes_factor:
        return a, None
    fac = np.frombuffer(a[:2].astype("<u2").tobytes(), dtype="<f4")[0]
    if fac < 0:
        raise ValueError(f"Negative factor: {fac}")
    return a[2:], np.float32(fac) / np.float32(65535.0)


def _read_grid_dir(dir, includes_factor):
    labels, spectra, factors = [], [], []
    for fn in sorted(os.listdir(dir)):
        labels.append(label_from_filename(fn))
        a, f = read_spectrum(_npy_u16(os.path.join(dir, fn), fn), includes_factor)
        spectra.append(a)
        factors.append(f)
    return labels, spectra, factors


def _read_grid_tar_zstd(file_path, includes_factor):
    """InMemFetcher::from_tar_zstd (src/model_fetchers.rs:168-225): a zstd-compressed tar
    stream of grid .npy files; members whose file name does not start with 'l' are
    skipped.  Decompressed as a stream (pyarrow's zstd codec), one member at a time."""
    import tarfile

    import pyarrow as pa

    labels, spectra, factors = [], [], []
    with pa.input_stream(file_path, compression="zstd") as raw:
        with tarfile.open(fileobj=raw, mode="r|") as tar:
            for m in tar:
                name = os.path.basename(m.name)
                if not m.isfile() or not name.startswith("l"):
                    continue
                labels.append(label_from_filename(name))
                a, f = read_spectrum(_npy_u16(tar.extractfile(m).read(), name), includes_factor)
                spectra.append(a)
                factors.append(f)
    if not labels:
        raise ValueError(f"no grid spectra in {file_path}")
    return labels, spectra, factors


class InMemInterpolator:
    """Grid resident in HBM (InMemInterpolator, src/lib.rs:1881-1908).

    InMemInterpolator(dir, includes_factor, wavelength) reads `l{m|p}MMMM_TTTTT_GGGG.npy`
    files like InMemFetcher::new (src/model_fetchers.rs:23-38,48-67,136-166);
    InMemInterpolator.from_arrays(spectra, labels, wavelength, factors) takes arrays.
    """

    def __init__(self, dir=None, includes_factor=False, wavelength=None, *, _arrays=None):
        if _arrays is None:
            labels, spectra, factors = _read_grid_dir(dir, includes_factor)
            spectra = np.stack(spectra)
            labels = np.array(labels)
            factors = np.array(factors, np.float32) if includes_factor else None
        else:
            spectra, labels, factors = _arrays
        if not wavelength.log:
            raise ValueError("only log-spaced synthetic grids are supported")
        spectra = np.ascontiguousarray(spectra, dtype=np.uint16)
        labels = f64(labels).reshape(-1, 3)
        if spectra.shape[1] != wavelength.len:
            raise ValueError("Length of model and wavelength grid don't match")
        self.wavelength = wavelength
        h = C.c_void_p()
        fac = None if factors is None else f32(factors)
        rc = lib.pasta_grid_create(ptr(spectra, C.c_uint16),
                                   None if fac is None else ptr(fac, C.c_float),
                                   ptr(labels, C.c_double), spectra.shape[0], spectra.shape[1],
                                   wavelength.min, wavelength.step, C.byref(h))
        _raise(rc, "grid")
        self._h = h
        shape = (C.c_int * 5)()
        lib.pasta_grid_shape(self._h, shape)
        self.n_teff, self.n_m, self.n_logg, self.n_models, self.n_pix = list(shape)

    @staticmethod
    def from_tar_zstd(file_path, includes_factor, wavelength):
        """The grid from a .tar.zst archive of .npy files (src/lib.rs:1896-1907)."""
        labels, spectra, factors = _read_grid_tar_zstd(file_path, includes_factor)
        n = {len(a) for a in spectra}
        if len(n) != 1:
            raise ValueError("grid spectra of different lengths in the archive")
        return InMemInterpolator(wavelength=wavelength, _arrays=(
            np.stack(spectra), np.array(labels),
            np.array(factors, np.float32) if includes_factor else None))

    @staticmethod
    def from_arrays(spectra, labels, wavelength, factors=None):
        return InMemInterpolator(wavelength=wavelength, _arrays=(spectra, labels, factors))

    @staticmethod
    def empty(labels, wavelength):
        """Allocate the grid in HBM without data (filled in place by a device producer)."""
        obj = InMemInterpolator.__new__(InMemInterpolator)
        labels = f64(labels).reshape(-1, 3)
        h = C.c_void_p()
        _raise(lib.pasta_grid_create(None, None, ptr(labels, C.c_double), len(labels),
                                     wavelength.len, wavelength.min, wavelength.step, C.byref(h)),
               "grid")
        obj._h, obj.wavelength = h, wavelength
        shape = (C.c_int * 5)()
        lib.pasta_grid_shape(h, shape)
        obj.n_teff, obj.n_m, obj.n_logg, obj.n_models, obj.n_pix = list(shape)
        return obj

    def __del__(self):
        if getattr(self, "_h", None):
            lib.pasta_grid_destroy(self._h)
            self._h = None

    # -- geometry accessors (src/lib.rs:1435-1465)
    def _axes(self):
        t = np.zeros(self.n_teff); m = np.zeros(self.n_m); g = np.zeros(self.n_logg)
        lo = np.zeros(self.n_teff, np.int32); hi = np.zeros(self.n_teff, np.int32)
        cum = np.zeros(self.n_teff, np.int32)
        lib.pasta_grid_axes(self._h, ptr(t, C.c_double), ptr(m, C.c_double), ptr(g, C.c_double),
                            ptr(lo, C.c_int), ptr(hi, C.c_int), ptr(cum, C.c_int))
        return t, m, g, lo, hi, cum

    def teffs(self):
        return self._axes()[0].tolist()

    def ms(self):
        return self._axes()[1].tolist()

    def loggs(self):
        return self._axes()[2].tolist()

    def logg_limits(self):
        _, _, _, lo, hi, _ = self._axes()
        return list(zip(lo.tolist(), hi.tolist()))

    def cumulative_grid_size(self):
        return self._axes()[5].tolist()

    def list_gridpoints(self):
        t, m, g, lo, hi, _ = self._axes()
        return [[t[i], mm, g[k]] for i in range(len(t)) for k in range(lo[i], hi[i] + 1) for mm in m]

    def is_teff_logg_between_bounds(self, teff, logg):
        return bool(lib.pasta_grid_is_teff_logg_between_bounds(self._h, teff, logg))

    def set_fft_taps(self, min_taps=0):
        """Rotation kernels of at least `min_taps` taps are convolved by FFT (where the
        tiles fit one 4096-point transform), shorter ones directly on the matrix cores; 0
        restores the measured default, a value above 2048 turns the FFT path off.  The
        reference always convolves by FFT (src/convolve_rv.rs:484-529)."""
        _raise(lib.pasta_grid_set_fft_taps(self._h, int(min_taps)), "set_fft_taps")

    def clamp_1d(self, teff, m, logg, index):
        out = C.c_double()
        _raise(lib.pasta_grid_clamp_1d(self._h, teff, m, logg, index, C.byref(out)), "clamp_1d")
        return out.value

    def _weights(self, teff, m, logg):
        idx = np.zeros(64, np.int32)
        w = np.zeros(64, np.float32)
        _raise(lib.pasta_grid_local_weights(self._h, teff, m, logg, ptr(idx, C.c_int),
                                            ptr(w, C.c_float)), "local grid")
        return idx, w

    def calculate_interpolation_coefficients(self, teff, m, logg):
        return self._weights(teff, m, logg)[1].tolist()

    def get_neighbor_indices(self, teff, m, logg):
        """(teff index, m index, logg index) per neighbour, (0,0,0) if absent (src/lib.rs:1414-1433)."""
        idx, _ = self._weights(teff, m, logg)
        t, mm, g, lo, hi, cum = self._axes()
        out = []
        for r in idx:
            if r < 0:
                out.append((0, 0, 0))
                continue
            j = r % len(mm)
            pair = r // len(mm)
            i = int(np.searchsorted(cum, pair, side="right") - 1)
            out.append((i, j, int(pair - cum[i] + lo[i])))
        return out

    # -- model production
    def interpolate(self, teff, m, logg):
        lab = f64([[teff, m, logg]])
        out = np.zeros(self.n_pix, np.float32)
        _raise(lib.pasta_interpolate(self._h, ptr(lab, C.c_double), 1, ptr(out, C.c_float), None),
               "Error while interpolating")
        return out

    def interpolate_bulk(self, labels):
        lab = f64(labels).reshape(-1, 3)
        out = np.zeros((len(lab), self.n_pix), np.float32)
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_interpolate(self._h, ptr(lab, C.c_double), len(lab),
                                        ptr(out, C.c_float), ptr(st, C.c_int)), "interpolate")
        return out, st

    def interpolate_linear(self, teff, m, logg):
        """Trilinear interpolation of the 8 neighbours (src/lib.rs:1127-1139)."""
        out, st = self.interpolate_linear_bulk([[teff, m, logg]])
        if st[0]:
            raise ValueError(f"Error while interpolating: {_lib.STATUS_TEXT.get(int(st[0]))}")
        return out[0]

    def interpolate_linear_bulk(self, labels):
        lab = f64(labels).reshape(-1, 3)
        out = np.zeros((len(lab), self.n_pix), np.float32)
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_interpolate_linear(self._h, ptr(lab, C.c_double), len(lab),
                                               ptr(out, C.c_float), ptr(st, C.c_int)),
                  "interpolate_linear")
        return out, st

    def produce_binary_model_norm(self, continuum_interpolator, dispersion, labels1, labels2,
                                  light_ratio):
        """Normalised model of a binary (src/lib.rs:1013-1035): the two components weighted
        by the linearly interpolated continua of `continuum_interpolator`."""
        out, st = self.produce_binary_model_norm_bulk(
            continuum_interpolator, dispersion,
            [list(labels1) + list(labels2) + [light_ratio]])
        if st[0]:
            raise ValueError(f"produce_binary_model_norm: {_lib.STATUS_TEXT.get(int(st[0]))}")
        return out[0]

    def produce_binary_model_norm_bulk(self, continuum_interpolator, dispersion, labels11):
        """Many binary models (labels P x 11) in one batched evaluation; (models, status)."""
        b = _Binary(self, continuum_interpolator, dispersion, None, None, None)
        return b.model(labels11)

    def get_binary_fitter(self, dispersion, continuum_fitter, settings, vsini_range, rv_range):
        """BinaryFitter (src/lib.rs:1177-1193)."""
        return BinaryFitter(dispersion, continuum_fitter, settings, vsini_range, rv_range)

    def _check_wl(self, synth_wl):
        w = self.wavelength
        if (synth_wl.log != w.log or synth_wl.len != w.len or synth_wl.min != w.min
                or synth_wl.step != w.step):
            raise ValueError("synth_wl differs from the interpolator's wavelength grid")

    def get_binary_rv_fitter(self, dispersion, synth_wl, continuum_fitter, settings, rv_range):
        """BinaryRVFitter (src/lib.rs:1195-1210): fits (rv1, rv2) of fixed components."""
        self._check_wl(synth_wl)
        return BinaryRVFitter(self, dispersion, continuum_fitter, settings, rv_range)

    def get_timeseries_known_rv_fitter(self, dispersion, synth_wl, continuum_fitter, settings,
                                       vsini_range):
        """BinaryTimeriesKnownRVFitter (src/lib.rs:1212-1227)."""
        self._check_wl(synth_wl)
        return BinaryTimeseriesKnownRVFitter(dispersion, continuum_fitter, settings, vsini_range)

    def interpolate_and_convolve(self, dispersion, teff, m, logg, vsini):
        lab = f64([[teff, m, logg, vsini]])
        out = np.zeros(self.n_pix, np.float32)
        _raise(lib.pasta_interpolate_and_convolve(self._h, dispersion.handle(self),
                                                  ptr(lab, C.c_double), 1, ptr(out, C.c_float),
                                                  None), "interpolate_and_convolve")
        return out

    def produce_model(self, dispersion, teff, m, logg, vsini, rv):
        out, st = self._produce(dispersion, [[teff, m, logg, vsini, rv]])
        if st[0]:
            raise ValueError(f"produce_model: {_lib.STATUS_TEXT.get(int(st[0]))}")
        return out[0]

    def produce_model_bulk(self, dispersion, labels, progress=None):
        out, st = self._produce(dispersion, labels)
        if st.any():
            raise ValueError(f"produce_model: {_lib.STATUS_TEXT.get(int(st[st != 0][0]))}")
        return out

    def _produce(self, dispersion, labels):
        lab = f64(labels).reshape(-1, 5)
        n = len(dispersion.wl)
        out = np.zeros((len(lab), n), np.float32)
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_produce_model(self._h, dispersion.handle(self), ptr(lab, C.c_double),
                                          len(lab), ptr(out, C.c_float), ptr(st, C.c_int)),
                  "produce_model")
        return out, st

    def chi2_bulk(self, observed_wavelength, observed_flux, observed_var, labels, progress=None,
                  allow_nan=None):
        """chi2 of many spectra, one label each, with NoConvolution and ChunkFitter(5, 8, 0.2)
        as the reference hard-codes (src/lib.rs:1297-1338); one batched GPU launch.  A
        failed model gives NaN with allow_nan=True, else raises."""
        wlg = self.wavelength
        stars = [Star(self, NoConvolutionDispersion(wl, wlg), ChunkContinuumFitter(wl, 5, 8, 0.2),
                      f, v) for wl, f, v in zip(observed_wavelength, observed_flux, observed_var)]
        n = len(stars)
        lab = f64(labels).reshape(n, 5)
        out = np.zeros(n)
        st = np.zeros(n, np.int32)
        arr = (C.c_void_p * n)(*[s._h for s in stars])
        rc = lib.pasta_eval_chi2_multi(arr, n, ptr(lab, C.c_double), ptr(out, C.c_double),
                                       ptr(st, C.c_int))
        if rc and not (allow_nan and rc == _lib.PASTA_ERR_MODEL):
            _raise(rc, "chi2_bulk")
        return out.tolist()

    def get_fitter(self, dispersion, continuum_fitter, settings, vsini_range, rv_range):
        return InMemSingleFitter(self, dispersion, continuum_fitter, settings, vsini_range, rv_range)


def get_vsini_kernel(vsini, synth_wl):
    """Rotation kernel (src/lib.rs:1962-1969)."""
    if not synth_wl.log:
        raise ValueError("Only logspace is supported")
    tmp = InMemInterpolator.__new__(InMemInterpolator)
    # the kernel depends only on the log step: compute it with a 2-pixel dummy grid
    lab = f64([[1.0, 0.0, 0.0]])
    h = C.c_void_p()
    spectra = np.zeros((1, 2), np.uint16)
    _raise(lib.pasta_grid_create(ptr(spectra, C.c_uint16), None, ptr(lab, C.c_double), 1, 2,
                                 synth_wl.min, synth_wl.step, C.byref(h)), "grid")
    try:
        n = C.c_int()
        lib.pasta_vsini_kernel(h, vsini, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), np.float32)
        _raise(lib.pasta_vsini_kernel(h, vsini, ptr(out, C.c_float), len(out), C.byref(n)), "kernel")
        return out[: n.value].tolist()
    finally:
        lib.pasta_grid_destroy(h)
        del tmp


def _settings_c(settings, vsini_range, rv_range, constraints):
    s = _lib.FitSettings()
    s.num_particles = settings.num_particles
    s.max_iters = settings.max_iters
    s.inertia_factor = settings.inertia_factor
    s.cognitive_factor = settings.cognitive_factor
    s.social_factor = settings.social_factor
    s.delta = settings.delta
    s.vsini_range[0], s.vsini_range[1] = vsini_range
    s.rv_range[0], s.rv_range[1] = rv_range
    for c in constraints or []:
        s.constraint_kind[c.parameter] = c.kind
        s.constraint_a[c.parameter] = c.a
        s.constraint_b[c.parameter] = c.b
    return s


class Star:
    """An observed spectrum resident in HBM, bound to a grid/dispersion/continuum."""

    def __init__(self, interpolator, dispersion, continuum, flux, var, _own=True):
        flux, var = f32(flux), f32(var)
        if len(flux) != len(var):
            raise ValueError("Flux and variance vectors must have the same length")
        h = C.c_void_p()
        _raise(lib.pasta_star_create(interpolator._h, dispersion.handle(interpolator),
                                     continuum.handle(len(flux)), ptr(flux, C.c_float),
                                     ptr(var, C.c_float), len(flux), C.byref(h)), "star")
        self._h, self._own, self.n = h, _own, len(flux)
        self.n_params = continuum.n_params(len(flux))
        self._refs = (interpolator, dispersion, continuum)

    def release(self):
        """Hand ownership to a server (the server destroys it)."""
        self._own = False
        return self._h

    def __del__(self):
        if getattr(self, "_own", False) and self._h:
            lib.pasta_star_destroy(self._h)
            self._h = None

    def chi2(self, labels):
        lab = f64(labels).reshape(-1, 5)
        out = np.zeros(len(lab))
        st = np.zeros(len(lab), np.int32)
        params = np.zeros((len(lab), max(1, self.n_params)), np.float32)
        _batch_rc(lib.pasta_eval_chi2(self._h, ptr(lab, C.c_double), len(lab),
                                      ptr(out, C.c_double), ptr(st, C.c_int),
                                      ptr(params, C.c_float)), "chi2")
        return out, st, params


class InMemSingleFitter:
    """SingleFitter for the resident grid (src/lib.rs:1468-1628)."""

    def __init__(self, interpolator, dispersion, continuum_fitter, settings, vsini_range, rv_range):
        self.interpolator, self.dispersion = interpolator, dispersion
        self.continuum_fitter, self.settings = continuum_fitter, settings
        self.vsini_range, self.rv_range = tuple(vsini_range), tuple(rv_range)

    def _star(self, interpolator, flux, var):
        if len(flux) != len(self.dispersion.wl):
            raise ValueError("Observed spectrum and target dispersion must have the same length")
        return Star(interpolator, self.dispersion, self.continuum_fitter, flux, var)

    def fit(self, interpolator, observed_flux, observed_var, trace_directory=None,
            best_particle_file=None, parallelize=None, constraints=None, seed=0, init_pos=None):
        """PSO fit of one star (src/lib.rs:1478-1507).  `seed` selects the random
        stream (the reference draws it from entropy); `parallelize` is accepted
        for compatibility (all particles always run on the GPU)."""
        if trace_directory is not None and best_particle_file is not None:
            raise ValueError("Cannot specify both trace_directory and best_particle_file")
        star = self._star(interpolator, observed_flux, observed_var)
        s = _settings_c(self.settings, self.vsini_range, self.rv_range, constraints)
        res = _lib.FitResult()
        trace = np.zeros((self.settings.max_iters + 1, 6))
        ip = None if init_pos is None else ptr(f64(init_pos), C.c_double)
        keep = None if init_pos is None else f64(init_pos)
        if keep is not None:
            ip = ptr(keep, C.c_double)
        if trace_directory is None:
            rc = lib.pasta_fit(star._h, C.byref(s), int(seed), ip, C.byref(res),
                               ptr(trace, C.c_double))
        else:
            pop = np.zeros((self.settings.max_iters, self.settings.num_particles, 6))
            rc = lib.pasta_fit_observed(star._h, C.byref(s), int(seed), ip, C.byref(res),
                                        ptr(trace, C.c_double), ptr(pop, C.c_double))
        _raise(rc, "fit")
        out = OptimizationResult._from_c(res)
        # the reference's observers (src/fitting.rs:217-296) see the state after each PSO
        # iteration (argmin's observe_iter, iteration numbers from 0), not the initialisation
        if best_particle_file is not None:
            with open(best_particle_file, "w") as f:
                json.dump([{"position": t[1:].tolist(), "cost": t[0]} for t in trace[1:]], f)
        if trace_directory is not None:
            for it in range(self.settings.max_iters):
                with open(os.path.join(trace_directory, f"it_{it}.json"), "w") as f:
                    json.dump([{"position": p[:5].tolist(), "cost": p[5]} for p in pop[it]], f)
        out._trace = trace
        return out

    def fit_bulk(self, interpolator, observed_fluxes, observed_vars, constraints=None, seeds=None,
                 trace=False):
        """Many stars in one device-driven batched PSO (src/lib.rs:1511-1540)."""
        stars = [self._star(interpolator, f, v) for f, v in zip(observed_fluxes, observed_vars)]
        n = len(stars)
        s = _settings_c(self.settings, self.vsini_range, self.rv_range, constraints)
        seeds = np.arange(n, dtype=np.uint64) if seeds is None else np.ascontiguousarray(seeds, np.uint64)
        arr = (C.c_void_p * n)(*[st._h for st in stars])
        res = (_lib.FitResult * n)()
        tr = np.zeros((n, self.settings.max_iters + 1, 6)) if trace else None
        rc = lib.pasta_fit_bulk(arr, n, C.byref(s), ptr(seeds, C.c_uint64), res,
                                None if tr is None else ptr(tr, C.c_double))
        _raise(rc, "fit_bulk")
        out = [OptimizationResult._from_c(r) for r in res]
        if trace:
            for k, o in enumerate(out):
                o._trace = tr[k]
        return out

    def compute_uncertainty(self, interpolator, observed_flux, observed_var, spec_res, label,
                            search_radius=None):
        """Uncertainties with the chi2 landscape method (src/lib.rs:1545-1568)."""
        star = self._star(interpolator, observed_flux, observed_var)
        out = np.zeros(10)
        st = np.zeros(10, np.int32)
        sr = None if search_radius is None else f64(search_radius)
        _raise(lib.pasta_compute_uncertainty(star._h, float(spec_res), ptr(f64(label), C.c_double),
                                             None if sr is None else ptr(sr, C.c_double),
                                             ptr(out, C.c_double), ptr(st, C.c_int)),
               "compute_uncertainty")
        return LabelUncertainties(out, st)

    def compute_uncertainty_bulk(self, interpolator, observed_fluxes, observed_vars, spec_res,
                                 labels, search_radius=None):
        """compute_uncertainty for many spectra in one batched search (src/lib.rs:1570-1600)."""
        stars = [self._star(interpolator, f, v) for f, v in zip(observed_fluxes, observed_vars)]
        n = len(stars)
        lab = f64(labels).reshape(n, 5)
        out = np.zeros((n, 10))
        st = np.zeros((n, 10), np.int32)
        sr = None if search_radius is None else f64(search_radius)
        arr = (C.c_void_p * n)(*[s._h for s in stars])
        _raise(lib.pasta_compute_uncertainty_bulk(arr, n, float(spec_res), ptr(lab, C.c_double),
                                                  None if sr is None else ptr(sr, C.c_double),
                                                  ptr(out, C.c_double), ptr(st, C.c_int)),
               "compute_uncertainty_bulk")
        return [LabelUncertainties(out[k], st[k]) for k in range(n)]

    def chi2(self, interpolator, observed_flux, observed_var, labels, allow_nan=None):
        """chi2 at each label vector (src/lib.rs:1603-1627): NaN on failure unless allow_nan=False."""
        star = self._star(interpolator, observed_flux, observed_var)
        out, st, _ = star.chi2(labels)
        if allow_nan is False and st.any():
            raise ValueError(_lib.STATUS_TEXT.get(int(st[st != 0][0])))
        return out.tolist()


class _Binary:
    """An observed binary resident in HBM (pasta_binary_*)."""

    def __init__(self, interpolator, continuum_interpolator, dispersion, continuum, flux, var):
        h = C.c_void_p()
        n = len(dispersion.wl)
        self.n_params = 0
        if flux is not None:
            flux, var = f32(flux), f32(var)
            if len(flux) != len(var):
                raise ValueError("Flux and variance vectors must have the same length")
            if len(flux) != n:
                raise ValueError("Observed spectrum and target dispersion must have the same length")
            self.n_params = continuum.n_params(n)
        rc = lib.pasta_binary_create(interpolator._h, continuum_interpolator._h,
                                     dispersion.handle(interpolator),
                                     None if continuum is None else continuum.handle(n),
                                     None if flux is None else ptr(flux, C.c_float),
                                     None if var is None else ptr(var, C.c_float), n, C.byref(h))
        _raise(rc, "binary")
        self._h, self.n = h, n
        self._refs = (interpolator, continuum_interpolator, dispersion, continuum)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.pasta_binary_destroy(self._h)
            self._h = None

    def model(self, labels11):
        lab = f64(labels11).reshape(-1, 11)
        out = np.zeros((len(lab), self.n), np.float32)
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_binary_model(self._h, ptr(lab, C.c_double), len(lab),
                                         ptr(out, C.c_float), ptr(st, C.c_int)), "binary model")
        return out, st

    def chi2(self, labels11, params=False):
        lab = f64(labels11).reshape(-1, 11)
        out = np.zeros(len(lab))
        st = np.zeros(len(lab), np.int32)
        prm = np.zeros((len(lab), max(1, self.n_params)), np.float32) if params else None
        _batch_rc(lib.pasta_binary_eval_chi2(self._h, ptr(lab, C.c_double), len(lab),
                                             ptr(out, C.c_double), ptr(st, C.c_int),
                                             None if prm is None else ptr(prm, C.c_float)),
                  "binary chi2")
        return out, st, prm


class _BinaryTS:
    """S spectra of one binary with known RVs (pasta_binary_ts_*)."""

    def __init__(self, interpolator, continuum_interpolator, dispersion, continuum, fluxes, vars_,
                 rvs):
        n = len(dispersion.wl)
        fl = [f32(f) for f in fluxes]
        vr = [f32(v) for v in vars_]
        if len(fl) != len(vr) or len(fl) != len(rvs) or not fl:
            raise ValueError("one flux, variance and RV pair per observed spectrum")
        for f, v in zip(fl, vr):
            if len(f) != len(v):
                raise ValueError("Flux and variance vectors must have the same length")
            if len(f) != n:
                raise ValueError("Observed spectrum and target dispersion must have the same length")
        F, V = np.ascontiguousarray(np.stack(fl)), np.ascontiguousarray(np.stack(vr))
        R = f64(rvs).reshape(-1, 2)
        h = C.c_void_p()
        _raise(lib.pasta_binary_ts_create(interpolator._h, continuum_interpolator._h,
                                          dispersion.handle(interpolator), continuum.handle(n),
                                          ptr(F, C.c_float), ptr(V, C.c_float), ptr(R, C.c_double),
                                          len(R), n, C.byref(h)), "binary time series")
        self._h, self.S, self.n_params = h, len(R), continuum.n_params(n)
        self._refs = (interpolator, continuum_interpolator, dispersion, continuum)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.pasta_binary_destroy(self._h)
            self._h = None

    def chi2(self, labels9):
        lab = f64(labels9).reshape(-1, 9)
        out = np.zeros(len(lab))
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_binary_ts_eval_chi2(self._h, ptr(lab, C.c_double), len(lab),
                                                ptr(out, C.c_double), ptr(st, C.c_int)),
                  "time-series chi2")
        return out, st


class _BinaryRV:
    """Fixed binary components, only the RVs free (pasta_binary_rv_*)."""

    def __init__(self, interpolator, dispersion, continuum, model1, model2, continuum1, continuum2,
                 light_ratio, flux, var):
        flux, var = f32(flux), f32(var)
        n = len(dispersion.wl)
        if len(flux) != len(var):
            raise ValueError("Flux and variance vectors must have the same length")
        if len(flux) != n:
            raise ValueError("Observed spectrum and target dispersion must have the same length")
        comps = [f32(a) for a in (model1, model2, continuum1, continuum2)]
        for a in comps:
            if len(a) != interpolator.n_pix:
                raise ValueError("Length of input_array and wavelength grid don't match")
        h = C.c_void_p()
        _raise(lib.pasta_binary_rv_create(interpolator._h, dispersion.handle(interpolator),
                                          continuum.handle(n), ptr(flux, C.c_float),
                                          ptr(var, C.c_float), n,
                                          *[ptr(a, C.c_float) for a in comps], float(light_ratio),
                                          C.byref(h)), "binary RV")
        self._h = h
        self._refs = (interpolator, dispersion, continuum)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.pasta_binary_rv_destroy(self._h)
            self._h = None

    def chi2(self, rvs):
        lab = f64(rvs).reshape(-1, 2)
        out = np.zeros(len(lab))
        st = np.zeros(len(lab), np.int32)
        _batch_rc(lib.pasta_binary_rv_eval_chi2(self._h, ptr(lab, C.c_double), len(lab),
                                                ptr(out, C.c_double), ptr(st, C.c_int)), "RV chi2")
        return out, st


def _binary_settings_c(settings, vsini_range, rv_range, constraints, n_labels=11):
    s = _lib.BinarySettings()
    s.num_particles = settings.num_particles
    s.max_iters = settings.max_iters
    s.inertia_factor = settings.inertia_factor
    s.cognitive_factor = settings.cognitive_factor
    s.social_factor = settings.social_factor
    s.delta = settings.delta
    s.vsini_range[0], s.vsini_range[1] = vsini_range
    s.rv_range[0], s.rv_range[1] = rv_range
    s.light_ratio_range[0], s.light_ratio_range[1] = 0.0, 1.0
    for c in constraints or []:
        if not 0 <= c.parameter < n_labels:
            raise ValueError(f"constraint parameter {c.parameter} outside the {n_labels} labels")
        s.constraint_kind[c.parameter] = c.kind
        s.constraint_a[c.parameter] = c.a
        s.constraint_b[c.parameter] = c.b
    return s


class BinaryFitter:
    """BinaryFitter (src/fitting.rs:606-728, Python src/lib.rs:1631-1700): PSO over the
    11 labels (star 1, star 2, light ratio) with BinaryBounds."""

    def __init__(self, dispersion, continuum_fitter, settings, vsini_range, rv_range):
        self.dispersion, self.continuum_fitter, self.settings = dispersion, continuum_fitter, settings
        self.vsini_range, self.rv_range = tuple(vsini_range), tuple(rv_range)

    def _binary(self, interpolator, continuum_interpolator, flux, var):
        return _Binary(interpolator, continuum_interpolator, self.dispersion,
                       self.continuum_fitter, flux, var)

    def fit(self, interpolator, continuum_interpolator, observed_flux, observed_var,
            trace_directory=None, best_particle_file=None, parallelize=None, constraints=None,
            seed=0, init_pos=None):
        """PSO fit of a binary (src/lib.rs:1637-1665); `seed` selects the random stream."""
        if trace_directory is not None and best_particle_file is not None:
            raise ValueError("Cannot specify both trace_directory and best_particle_file")
        if trace_directory is not None:
            raise NotImplementedError("the per-iteration population files (trace_directory) are "
                                      "written by single-star fits only")
        b = self._binary(interpolator, continuum_interpolator, observed_flux, observed_var)
        s = _binary_settings_c(self.settings, self.vsini_range, self.rv_range, constraints)
        res = _lib.BinaryResult()
        trace = np.zeros((self.settings.max_iters + 1, 12))
        keep = None if init_pos is None else f64(init_pos)
        rc = lib.pasta_binary_fit(b._h, C.byref(s), int(seed),
                                  None if keep is None else ptr(keep, C.c_double), C.byref(res),
                                  ptr(trace, C.c_double))
        _raise(rc, "fit")
        out = BinaryOptimizationResult._from_c(res)
        if best_particle_file is not None:
            with open(best_particle_file, "w") as f:
                json.dump([{"position": t[1:].tolist(), "cost": t[0]} for t in trace[1:]], f)
        out._trace = trace
        return out

    def chi2(self, interpolator, continuum_interpolator, observed_flux, observed_var, labels,
             allow_nan=None):
        """chi2 at each 11-label vector (src/lib.rs:1667-1699): NaN on failure unless
        allow_nan=False."""
        b = self._binary(interpolator, continuum_interpolator, observed_flux, observed_var)
        out, st, _ = b.chi2(labels)
        if allow_nan is False and st.any():
            raise ValueError(_lib.STATUS_TEXT.get(int(st[st != 0][0])))
        return out.tolist()


class BinaryRVFitter:
    """BinaryRVFitter (src/fitting.rs:785-880, Python src/lib.rs:1702-1779): PSO over
    (rv1, rv2) of fixed components (synthetic-grid models and continua)."""

    def __init__(self, interpolator, dispersion, continuum_fitter, settings, rv_range):
        self.interpolator, self.dispersion = interpolator, dispersion
        self.continuum_fitter, self.settings, self.rv_range = continuum_fitter, settings, tuple(rv_range)

    def _rv(self, model1, model2, continuum1, continuum2, light_ratio, flux, var):
        return _BinaryRV(self.interpolator, self.dispersion, self.continuum_fitter, model1, model2,
                         continuum1, continuum2, light_ratio, flux, var)

    def fit(self, model1, model2, continuum1, continuum2, light_ratio, observed_flux, observed_var,
            trace_directory=None, best_particle_file=None, constraints=None, seed=0, init_pos=None):
        if trace_directory is not None and best_particle_file is not None:
            raise ValueError("Cannot specify both trace_directory and best_particle_file")
        if trace_directory is not None:
            raise NotImplementedError("the per-iteration population files (trace_directory) are "
                                      "written by single-star fits only")
        r = self._rv(model1, model2, continuum1, continuum2, light_ratio, observed_flux, observed_var)
        s = _binary_settings_c(self.settings, (0.0, 0.0), self.rv_range, constraints, 2)
        res = _lib.RvResult()
        trace = np.zeros((self.settings.max_iters + 1, 3))
        keep = None if init_pos is None else f64(init_pos)
        _raise(lib.pasta_binary_rv_fit(r._h, C.byref(s), int(seed),
                                       None if keep is None else ptr(keep, C.c_double),
                                       C.byref(res), ptr(trace, C.c_double)), "fit")
        out = RVOptimizationResult(res.rv1, res.rv2, res.chi2, res.iterations, res.time_s)
        if best_particle_file is not None:
            with open(best_particle_file, "w") as f:
                json.dump([{"position": t[1:].tolist(), "cost": t[0]} for t in trace[1:]], f)
        out._trace = trace
        return out

    def chi2(self, model1, model2, continuum1, continuum2, light_ratio, observed_flux, observed_var,
             labels, allow_nan=None):
        r = self._rv(model1, model2, continuum1, continuum2, light_ratio, observed_flux, observed_var)
        out, st = r.chi2(labels)
        if allow_nan is False and st.any():
            raise ValueError(_lib.STATUS_TEXT.get(int(st[st != 0][0])))
        return out.tolist()


class BinaryTimeseriesKnownRVFitter:
    """BinaryTimeriesKnownRVFitter (src/fitting.rs:987-1078, Python src/lib.rs:1781-1855):
    PSO over (Teff, [M/H], logg, vsini) of both stars and the light ratio, the chi2 summed
    over spectra with known RVs."""

    def __init__(self, dispersion, continuum_fitter, settings, vsini_range):
        self.dispersion, self.continuum_fitter = dispersion, continuum_fitter
        self.settings, self.vsini_range = settings, tuple(vsini_range)

    def _ts(self, interpolator, continuum_interpolator, fluxes, vars_, rvs):
        return _BinaryTS(interpolator, continuum_interpolator, self.dispersion,
                         self.continuum_fitter, fluxes, vars_, rvs)

    def fit(self, interpolator, continuum_interpolator, observed_fluxes, observed_vars, rvs,
            trace_directory=None, best_particle_file=None, parallelize=None, constraints=None,
            seed=0, init_pos=None):
        if trace_directory is not None and best_particle_file is not None:
            raise ValueError("Cannot specify both trace_directory and best_particle_file")
        if trace_directory is not None:
            raise NotImplementedError("the per-iteration population files (trace_directory) are "
                                      "written by single-star fits only")
        t = self._ts(interpolator, continuum_interpolator, observed_fluxes, observed_vars, rvs)
        s = _binary_settings_c(self.settings, self.vsini_range, (0.0, 0.0), constraints, 9)
        res = _lib.TsResult()
        np_ = max(1, t.n_params)
        prm = np.zeros((t.S, np_), np.float32)
        chis = np.zeros(t.S)
        trace = np.zeros((self.settings.max_iters + 1, 10))
        keep = None if init_pos is None else f64(init_pos)
        _raise(lib.pasta_binary_ts_fit(t._h, C.byref(s), int(seed),
                                       None if keep is None else ptr(keep, C.c_double),
                                       C.byref(res), ptr(prm, C.c_float), ptr(chis, C.c_double),
                                       ptr(trace, C.c_double)), "fit")
        out = BinaryTimeseriesOptimizationResult(
            RvLessLabel(*res.label1[:4]), RvLessLabel(*res.label2[:4]), res.light_ratio,
            [p[: t.n_params] for p in prm], chis, res.iterations, res.time_s)
        out.chi2 = res.chi2
        if best_particle_file is not None:
            with open(best_particle_file, "w") as f:
                json.dump([{"position": tr[1:].tolist(), "cost": tr[0]} for tr in trace], f)
        out._trace = trace
        return out

    def chi2(self, interpolator, continuum_interpolator, observed_fluxes, observed_vars, rvs, labels,
             allow_nan=None):
        t = self._ts(interpolator, continuum_interpolator, observed_fluxes, observed_vars, rvs)
        out, st = t.chi2(labels)
        if allow_nan is False and st.any():
            raise ValueError(_lib.STATUS_TEXT.get(int(st[st != 0][0])))
        return out.tolist()


class FitServer:
    """Continuous-batching fit server over a resident grid (pasta_server_*).

    Requests are admitted between steps while slots and HBM allow; every admitted
    star's particles run in one captured PSO-iteration graph; finished stars are
    retired at once.  A result is identical alone and in any mix of requests."""

    def __init__(self, interpolator, max_stars=512, max_particles=64, iters_per_step=1,
                 hbm_budget_bytes=0):
        cfg = _lib.ServerConfig(int(max_stars), int(max_particles), int(iters_per_step),
                                int(hbm_budget_bytes), 0)
        h = C.c_void_p()
        _raise(lib.pasta_server_create(interpolator._h, C.byref(cfg), C.byref(h)), "server")
        self._h, self._interp = h, interpolator

    def submit(self, star, settings, vsini_range, rv_range, seed=0, constraints=None):
        """Queue a fit of `star` (a Star; the server takes it over).  Returns the request id."""
        s = _settings_c(settings, vsini_range, rv_range, constraints)
        rid = C.c_uint64()
        _raise(lib.pasta_server_submit(self._h, star._h, C.byref(s), int(seed), C.byref(rid)),
               "submit")
        star.release()
        return rid.value

    def step(self):
        """Admit, run iters_per_step PSO iterations of every admitted star, retire.
        Returns (active stars, waiting requests)."""
        a, w = C.c_int(), C.c_int()
        _raise(lib.pasta_server_step(self._h, C.byref(a), C.byref(w)), "step")
        return a.value, w.value

    def poll(self, request_id):
        """OptimizationResult of a finished request, else None."""
        done = C.c_int()
        r = _lib.FitResult()
        _raise(lib.pasta_server_poll(self._h, int(request_id), C.byref(done), C.byref(r)), "poll")
        return OptimizationResult._from_c(r) if done.value else None

    def cancel(self, request_id):
        _raise(lib.pasta_server_cancel(self._h, int(request_id)), "cancel")

    def stats(self):
        """(evaluations, requests finished, iterations, HBM bytes in use, device-busy
        microseconds, batch rebuilds, host microseconds in rebuilds)."""
        out = (C.c_int64 * 7)()
        _raise(lib.pasta_server_stats(self._h, out), "stats")
        return tuple(out)

    def close(self):
        if getattr(self, "_h", None):
            lib.pasta_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
