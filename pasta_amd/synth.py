"""Seeded synthetic workloads (grids, observed spectra, batches).

Grid recipe: normalised flux prod_l (1 - d_l(u) exp(-(lambda-lambda_l)^2 / (2 w_l^2))),
lambda_l ~ U(lambda_min, lambda_max), w_l ~ U(0.05, 0.2) A,
d_l = sigmoid(a_l . u + b_l) with u the labels scaled to [0, 1]^3,
a_l ~ N(0, 2^2 I), b_l ~ N(-1, 1); stored as uint16(flux * 65535) (truncation,
as the reference's README prescribes for normalised grids).  The synthetic
wavelength grid is WlGrid::Logspace(log10 lambda_min, log10(lambda_max/lambda_min)/n, n).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class GridSpec:
    teff: np.ndarray
    m: np.ndarray
    logg: np.ndarray
    n_pix: int
    wl_min: float
    wl_max: float
    n_lines: int
    line_seed: int = 1000

    @property
    def log10_first(self):
        return float(np.log10(self.wl_min))

    @property
    def log10_step(self):
        return float(np.log10(self.wl_max / self.wl_min) / self.n_pix)

    def synth_wl(self):
        return 10.0 ** (self.log10_first + self.log10_step * np.arange(self.n_pix))

    def labels(self):
        """Grid points in storage order (teff-major, then logg, then [M/H])."""
        out = []
        for t in self.teff:
            for lg in self.logg:
                for mh in self.m:
                    out.append((t, mh, lg))
        return np.array(out, dtype=np.float64)

    def lines(self):
        rng = np.random.default_rng(self.line_seed)
        centre = rng.uniform(self.wl_min, self.wl_max, self.n_lines)
        width = rng.uniform(0.05, 0.2, self.n_lines)
        a = rng.normal(0.0, 2.0, (self.n_lines, 3))
        b = rng.normal(-1.0, 1.0, self.n_lines)
        order = np.argsort(centre, kind="stable")
        return centre[order], width[order], a[order], b[order]

    def unit_labels(self, labels):
        lo = np.array([self.teff.min(), self.m.min(), self.logg.min()])
        hi = np.array([self.teff.max(), self.m.max(), self.logg.max()])
        return (labels - lo) / (hi - lo)


def small_grid_spec():
    """Teff 5000-7000/250, [M/H] -1..0.5/0.25, logg 3.5-5/0.5; 20k px, 5000-5500 A, 200 lines."""
    return GridSpec(np.arange(5000.0, 7000.1, 250.0), np.round(np.arange(-1.0, 0.51, 0.25), 2),
                    np.arange(3.5, 5.01, 0.5), 20_000, 5000.0, 5500.0, 200)


def full_grid_spec():
    """Teff 4000-10000/250, [M/H] -1..0.5/0.1, logg 3-5/0.1; 200k px, 3800-9000 A, 5000 lines."""
    return GridSpec(np.arange(4000.0, 10000.1, 250.0), np.round(np.arange(-1.0, 0.501, 0.1), 2),
                    np.round(np.arange(3.0, 5.001, 0.1), 2), 200_000, 3800.0, 9000.0, 5000)


def make_grid(spec, labels=None, includes_factor_seed=None):
    """uint16 spectra (n_models x n_pix) for `labels` (default: every grid point, storage
    order).  With includes_factor_seed, each model is multiplied by a continuum whose
    level rises smoothly with Teff from 1e3 to 1e7, stored as flux/max, and the factors
    (max/65535, f32) are returned too."""
    if labels is None:
        labels = spec.labels()
    wl = spec.synth_wl()
    centre, width, a, b = spec.lines()
    u = spec.unit_labels(labels)
    depth = 1.0 / (1.0 + np.exp(-(u @ a.T + b[None, :])))  # (models, lines)
    flux = np.ones((len(labels), spec.n_pix), dtype=np.float64)
    lo = np.searchsorted(wl, centre - 8 * width)
    hi = np.searchsorted(wl, centre + 8 * width)
    for l in range(spec.n_lines):
        s, e = lo[l], hi[l]
        if e <= s:
            continue
        prof = np.exp(-0.5 * ((wl[s:e] - centre[l]) / width[l]) ** 2)
        flux[:, s:e] *= 1.0 - depth[:, l:l + 1] * prof[None, :]
    if includes_factor_seed is None:
        return (flux * 65535.0).astype(np.uint16), None
    rng = np.random.default_rng(includes_factor_seed)
    # continuum level rising smoothly with Teff over 1e3..1e7 (0.02 dex jitter)
    mx = 10.0 ** (3.0 + 4.0 * u[:, 0] + rng.normal(0.0, 0.02, len(labels)))
    x = np.linspace(-1, 1, spec.n_pix)
    shape = 1.0 + 0.3 * x[None, :] * rng.uniform(-1, 1, (len(labels), 1))
    cont = flux * shape
    cmax = cont.max(axis=1)
    u16 = (cont / cmax[:, None] * 65535.0).astype(np.uint16)
    factors = (np.float32(1.0) * (mx * cmax).astype(np.float32)) / np.float32(65535.0)
    return u16, factors.astype(np.float32)


def make_continuum_grid(spec, labels=None):
    """Unnormalised continuum grid for binary models: the Planck curve
    B(lambda, Teff) ~ lambda^-5 / (exp(hc / (lambda k Teff)) - 1) times 10^(0.2 [M/H]) and
    (1 + 0.05 (logg - 4)), stored like an includes_factor grid (uint16 flux / max and the
    f32 factor max / 65535 per model)."""
    if labels is None:
        labels = spec.labels()
    wl = spec.synth_wl()
    hc_k = 1.4387769e8  # Angstrom K
    # the stored rows depend on Teff only (flux / max); [M/H] and logg scale the factors
    teffs, ti = np.unique(labels[:, 0], return_inverse=True)
    rows = np.empty((len(teffs), len(wl)), np.uint16)
    pmax = np.empty(len(teffs))
    for k, t in enumerate(teffs):
        b = wl ** -5.0 / np.expm1(hc_k / (wl * t))
        pmax[k] = b.max()
        rows[k] = (b / pmax[k] * 65535.0).astype(np.uint16)
    scale = pmax[ti] * 10.0 ** (0.2 * labels[:, 1]) * (1.0 + 0.05 * (labels[:, 2] - 4.0))
    bmax = scale / scale.max() * 1e6
    factors = (bmax.astype(np.float32) / np.float32(65535.0)).astype(np.float32)
    return rows[ti], factors


def observed_wavelengths(spec, n_obs, margin_px):
    """Log-uniform observed sampling kept margin_px synthetic pixels inside each edge."""
    lo = spec.log10_first + spec.log10_step * margin_px
    hi = spec.log10_first + spec.log10_step * (spec.n_pix - margin_px)
    return 10.0 ** np.linspace(lo, hi, n_obs)


def observe(model_flux, wl, snr, seed):
    """Multiply by a seeded degree-3 continuum of scale 1e4 and add Gaussian noise."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.1, 0.1, 3)
    x = np.linspace(-1.0, 1.0, len(wl))
    cont = 1e4 * (1.0 + c[0] * x + c[1] * x ** 2 + c[2] * x ** 3)
    flux = model_flux.astype(np.float64) * cont
    var = (flux / snr) ** 2
    flux = flux + rng.normal(0.0, 1.0, len(wl)) * np.sqrt(var)
    return flux.astype(np.float32), var.astype(np.float32)


def random_labels(spec, n, seed, vsini=(0.0, 300.0), rv=(-150.0, 150.0), margin=0.0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(spec.teff.min() + margin, spec.teff.max() - margin, n)
    m = rng.uniform(spec.m.min(), spec.m.max(), n)
    g = rng.uniform(spec.logg.min(), spec.logg.max(), n)
    return np.stack([t, m, g, rng.uniform(*vsini, n), rng.uniform(*rv, n)], axis=1)


def smooth_design(n, ncols, seed):
    """Linear continuum design matrix (n x ncols): a constant column and seeded-phase
    cosines of k = 1 .. ncols-1 half periods over the range, so the columns are close to
    orthogonal (condition number ~7 at 12 columns, ~18 at 30)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, n)
    cols = [np.ones(n)]
    for k in range(1, ncols):
        ph = rng.uniform(0, np.pi)
        cols.append(np.cos(np.pi * k * x + ph))
    return np.stack(cols, axis=1).astype(np.float32)


def fill_grid_gpu(interp, spec):
    """Fill a resident grid (InMemInterpolator.empty(spec.labels(), ...)) in place on the
    GPU with the same recipe as make_grid (libpasta_synth.so)."""
    import ctypes as C
    import os

    from . import _lib

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpasta_synth.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python pasta_amd/build.py`")
    sl = C.CDLL(path)
    rows = C.POINTER(C.c_uint16)()
    pitch = C.c_int64()
    _lib.check(_lib.lib.pasta_grid_device_rows(interp._h, C.byref(rows), C.byref(pitch)))
    centre, width, a, b = spec.lines()
    ul = np.ascontiguousarray(spec.unit_labels(spec.labels()), dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    sl.pasta_synth_fill_grid.argtypes = [C.POINTER(C.c_uint16), C.c_int64, C.c_int, C.c_int,
                                         C.c_double, C.c_double, dp, C.c_int, dp, dp, dp, dp]
    rc = sl.pasta_synth_fill_grid(rows, pitch.value, len(ul), spec.n_pix, spec.log10_first,
                                  spec.log10_step, ul.ctypes.data_as(dp), spec.n_lines,
                                  centre.ctypes.data_as(dp), width.ctypes.data_as(dp),
                                  a.ctypes.data_as(dp), b.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(f"pasta_synth_fill_grid failed ({rc})")


def copy_probe(nbytes):
    """One plain device copy of nbytes (for scaling FETCH_SIZE / WRITE_SIZE)."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpasta_synth.so")
    sl = C.CDLL(path)
    sl.pasta_synth_copy_probe.argtypes = [C.c_int64]
    return sl.pasta_synth_copy_probe(int(nbytes))


def read_grid(interp):
    """Resident grid rows back to the host (storage order)."""
    import ctypes as C
    from . import _lib
    out = np.zeros((interp.n_models, interp.n_pix), np.uint16)
    _lib.check(_lib.lib.pasta_grid_read_rows(interp._h, 0, interp.n_models,
                                             out.ctypes.data_as(C.POINTER(C.c_uint16))))
    return out
