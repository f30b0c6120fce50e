"""Pins the oracle's functions that the closed-form checks elsewhere leave open, each by
what the mathematics fixes rather than by retyping the oracle:

* the Doppler direction of produce_model (src/convolve_rv.rs:547,587): a single line at
  lambda_0 lands at lambda_0 / (1 - rv/299792) in the observed frame;
* the "same"-mode offsets of the rotation (start = K/2, src/convolve_rv.rs:523) and of the
  fixed-resolution Gaussian (rows((K-1)/2), :209): both stages equal numpy.convolve(...,
  "same") with zero padding in FP64, and a symmetric line keeps its centroid;
* VariableTargetDispersion (:303-431): every kernel row is the unit-sum Gaussian of
  sigma_i = disp_i / (step ln10 lambda_i), K = ceil(6 max sigma) made odd, the per-pixel
  dot equals the FP64 sum and the first and last (K-1)/2 pixels are copied through;
* two PSO iterations of a 2-particle swarm computed by hand from the counter-RNG draws
  (src/particleswarm.rs:247-298): momentum, the cognitive and the social pull with
  distinct factors, the per-dimension clamp order (Teff with the old logg, logg with the
  new Teff, src/bounds.rs:13-16, src/interpolate.rs:647-664) and the strict "<" of the
  personal and global best updates;
* build_continuum_from_chunks (src/continuum_fitting.rs:228-278): the FP64 blend of the
  chunk polynomials, and the host library's pasta_cont_build equal to the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

C_KMS = 299792.0
LN10 = np.log(10.0)

TEFF4 = np.array([5000.0, 5250.0, 5500.0, 5750.0])
MH4 = np.array([-0.5, -0.25, 0.0, 0.25])
LOGG4 = np.array([4.0, 4.25, 4.5, 4.75])
W = 8192  # a multiple of 128: no interpolation tail
WL0, WL1 = 5000.0, 5100.0


def _grid_labels():
    return np.array([(t, m, g) for t in TEFF4 for g in LOGG4 for m in MH4])


def _flat_grid(flux):
    """A 4 x 4 x 4 grid whose every model is `flux` (W pixels, log-uniform 5000-5100 A):
    interpolation returns `flux` quantised to uint16 at any label."""
    step = np.log10(WL1 / WL0) / W
    u16 = np.floor(np.clip(flux, 0, 1) * 65535.0).astype(np.uint16)
    spectra = np.repeat(u16[None, :], 64, axis=0)
    return O.Grid(_grid_labels(), spectra, None, np.log10(WL0), step), step


def _synth_wl(step):
    return 10.0 ** (np.log10(WL0) + step * np.arange(W))


def _line(wl, centre, depth, sigma_a):
    return 1.0 - depth * np.exp(-0.5 * ((wl - centre) / sigma_a) ** 2)


def _centroid_px(flux, u):
    a = 1.0 - np.asarray(flux, np.float64)
    return float((a * u).sum() / a.sum())


@pytest.mark.parametrize("rv", [100.0, -100.0, 37.5, 0.0])
def test_doppler_direction(rv):
    step = np.log10(WL1 / WL0) / W
    wl_s = _synth_wl(step)
    lam0 = 5050.0
    og, step = _flat_grid(_line(wl_s, lam0, 0.6, 0.05))
    # observed pixels: the synthetic sampling, 400 pixels inside each edge
    wl = wl_s[400:-400]
    od = O.Dispersion(og, O.DISP_NONE, wl)
    u = (np.log10(wl) - np.log10(WL0)) / step
    for f64 in (False, True):
        m = O.produce_model(og, od, [5300.0, -0.1, 4.3, 0.0, rv], f64=f64)
        win = np.abs(u - (np.log10(lam0) - np.log10(WL0)) / step) < 150
        got = _centroid_px(m[win], u[win])
        want = (np.log10(lam0 / (1.0 - rv / C_KMS)) - np.log10(WL0)) / step
        # a sign error would move the line by 2 rv / c / (ln10 step) = 70 px at 100 km/s
        assert abs(got - want) < 0.05, (rv, f64, got, want)


def _line_rich(wl, seed, n=60):
    rng = np.random.default_rng(seed)
    f = np.ones_like(wl)
    for c, d, s in zip(rng.uniform(WL0, WL1, n), rng.uniform(0.1, 0.8, n),
                       rng.uniform(0.02, 0.15, n)):
        f *= _line(wl, c, d, s)
    return f


@pytest.mark.parametrize("vsini", [8.0, 60.0, 250.0])
def test_same_mode_offsets_and_zero_padding(vsini):
    step = np.log10(WL1 / WL0) / W
    wl_s = _synth_wl(step)
    flux = _line_rich(wl_s, 3)
    flux[:30] = 0.7  # edges well below 1: zero padding is visible there
    og, step = _flat_grid(flux)
    wl = wl_s[500:-500]
    R = 40_000.0
    od = O.Dispersion(og, O.DISP_FIXED, wl, resolution=R)
    lab = [5300.0, -0.1, 4.3, vsini]
    got = O.interpolate_and_convolve(og, od, lab).astype(np.float64)
    interp = og.interpolate(*lab[:3]).astype(np.float64)
    rot = O.rotation_kernel(vsini, LN10 * step).astype(np.float64)
    gauss = od.kernels(W).astype(np.float64)
    assert len(rot) % 2 == 1 and len(gauss) % 2 == 1
    want = np.convolve(np.convolve(interp, rot, "same"), gauss, "same")
    tol = 2e-6 * np.abs(want).max()  # the reference's f32 FFT rounding
    np.testing.assert_allclose(got, want, rtol=0, atol=tol)
    # an offset one pixel off, or edge copy-through instead of zero padding, fails
    assert np.abs(got - np.roll(want, 1)).max() > 100 * tol
    copy = want.copy()
    h = (len(rot) - 1) // 2
    copy[:h] = interp[:h]
    assert np.abs(got - copy).max() > 100 * tol


def test_symmetric_line_keeps_its_centroid():
    step = np.log10(WL1 / WL0) / W
    wl_s = _synth_wl(step)
    lam0 = wl_s[W // 2]
    og, step = _flat_grid(_line(wl_s, lam0, 0.5, 0.04))
    od = O.Dispersion(og, O.DISP_FIXED, wl_s[500:-500], resolution=50_000.0)
    u = np.arange(W, dtype=np.float64)
    hg = (od.K - 1) // 2
    for vsini in (20.0, 120.0):
        out = O.interpolate_and_convolve(og, od, [5300.0, -0.1, 4.3, vsini])
        # over the broadened line's support (line +- 5 sigma, both kernels' half-widths):
        # outside it the f32 FFT leaves ~2e-7 of continuum offset, which a wide window
        # would turn into a centroid bias
        hr = (len(O.rotation_kernel(vsini, LN10 * step)) - 1) // 2
        win = np.abs(u - W // 2) <= hr + hg + 20
        assert abs(_centroid_px(out[win], u[win]) - W // 2) < 0.01, vsini


def _resample_to_synth(wl_obs, wl_s, y):
    """resample_observed_to_synth (src/convolve_rv.rs:303-331) in FP64: linear
    interpolation in wavelength, the last observed value beyond the end."""
    out = np.empty(len(wl_s))
    j = 0
    for i, x in enumerate(wl_s):
        while j < len(wl_obs) - 1 and wl_obs[j + 1] < x:
            j += 1
        if j == len(wl_obs) - 1:
            out[i] = y[j]
        else:
            w = (x - wl_obs[j]) / (wl_obs[j + 1] - wl_obs[j])
            out[i] = (1 - w) * y[j] + w * y[j + 1]
    return out


def test_variable_dispersion_kernels_and_edges():
    step = np.log10(WL1 / WL0) / W
    wl_s = _synth_wl(step)
    og, step = _flat_grid(_line_rich(wl_s, 5))
    wl = 10.0 ** np.linspace(np.log10(WL0) + 300 * step, np.log10(WL1) - 300 * step, 5000)
    R = np.interp(wl, [wl[0], wl[-1]], [30_000.0, 60_000.0])
    disp = (wl / R).astype(np.float32)  # sigma in A: the code divides by no 2.355
    od = O.Dispersion(og, O.DISP_VARIABLE, wl, disp=disp)
    K = od.K
    sig = _resample_to_synth(wl, wl_s, disp.astype(np.float64)) / (step * LN10 * wl_s)
    kme = int(np.ceil(6 * sig.max()))
    assert K == kme + 1 - kme % 2 and K % 2 == 1
    h = (K - 1) // 2
    ker = od.kernels(W).astype(np.float64)
    x = np.arange(K) - h
    ref = np.exp(-0.5 * (x[None, :] / sig[:, None]) ** 2)
    ref /= ref.sum(axis=1, keepdims=True)
    # inside the observed span the f32 resample of the dispersion is accurate to f32; before
    # the first observed pixel it extrapolates with weights of ~-200, and the reference's f32
    # arithmetic (1 - w) y0 + w y1 loses ~8 bits there
    inside = (wl_s >= wl[0]) & (wl_s <= wl[-1])
    np.testing.assert_allclose(ker[inside], ref[inside], rtol=2e-5, atol=1e-7)
    np.testing.assert_allclose(ker[~inside], ref[~inside], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(ker.sum(axis=1), 1.0, atol=2e-6)
    # per-pixel dot and edge copy-through
    lab = [5300.0, -0.1, 4.3, 0.0]
    got = O.interpolate_and_convolve(og, od, lab).astype(np.float64)
    flat = og.interpolate(*lab[:3]).astype(np.float64)
    want = flat.copy()
    win = np.lib.stride_tricks.sliding_window_view(flat, K)
    want[h:W - h] = (win * ker[h:W - h]).sum(axis=1)  # the kernels pinned above, FP64 dot
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-7)
    assert np.array_equal(got[:h], flat[:h]) and np.array_equal(got[W - h:], flat[W - h:])


# --------------------------------------------------------------------------- PSO
from tests.test_oracle_grid import LOGG, MH, STAIR, TEFF, staircase_labels  # noqa: E402

KEY_ITER = 0x51A5A0000000003


def _u(key, it, P, q, d, k):
    """Draw k (0: cognitive, 1: social) of particle q, dimension d, iteration it."""
    return O._L.or_rng_uniform(key, (((it * P + q) * 5 + d) * 2) + k)


def _rand_range(a, b, u):  # random_uniform (particleswarm.rs:153-161): nothing when a == b
    if a == b:
        return a
    return a + (b - a) * u if a < b else b + (a - b) * u


def test_pso_two_iterations_by_hand():
    """STAIR grid: Teff 5000-7000 / 250, logg 3-5 / 0.5, per-Teff logg index limits
    (0,4) x3, (1,4) x2, (1,3), (2,3) x3.  At logg 3.2 only Teff <= 5500 exist; at Teff
    5500 (a node whose right neighbour has limits (1,4)) logg spans 3.5-5.0."""
    g = O.Grid(staircase_labels(TEFF, MH, LOGG, STAIR))
    b = O.bounds(g, (0.0, 300.0), (-150.0, 150.0))
    inertia, cog, soc = 1.0, 0.7, 1.9
    s = O.pso_settings(2, 2, inertia, cog, soc)
    seed = 11
    key = O._L.or_rng_u64(seed, KEY_ITER)
    x0 = np.array([[5400.0, 0.0, 3.2, 100.0, 10.0],    # sorted first (lower cost)
                   [6000.0, -0.5, 4.2, 50.0, -20.0]])
    v0 = np.array([[900.0, 0.0, -0.1, 0.0, 0.0],
                   [-40.0, 0.05, 0.02, 5.0, -3.0]])
    seen = []
    costs = [np.array([1.0, 2.0]), np.array([0.5, 3.0]), np.array([0.5, 0.5])]

    def f(pos):
        seen.append(pos.copy())
        return costs[len(seen) - 1]

    best, bc, trace, ne = O.pso_run(b, s, seed, f, init_pos=x0, init_vel=v0)
    assert ne == 6 and len(seen) == 3
    np.testing.assert_array_equal(seen[0], x0)

    # iteration 1: pbest = own start, gbest = particle 0's start, so only momentum and
    # the social pull act.  Particle 0: Teff 5400 + 900 = 6300 is clamped with the OLD
    # logg 3.2 to 5500; logg 3.2 - 0.1 = 3.1 is clamped with the NEW Teff 5500 to 3.5.
    x1 = x0.copy(); v1 = v0.copy()
    gbest = x0[0]
    for q in range(2):
        for d in range(5):
            pull = _rand_range(0.0, x0[q, d] - x1[q, d], _u(key, 1, 2, q, d, 0)) * cog
            pullg = _rand_range(0.0, gbest[d] - x1[q, d], _u(key, 1, 2, q, d, 1)) * soc
            assert pull == 0.0
            v1[q, d] = v1[q, d] * inertia + pull + pullg
            x1[q, d] = x1[q, d] + v1[q, d]
    x1[0, 0] = 5500.0  # Teff limits at logg 3.2: 5000-5500
    x1[0, 2] = 3.5     # logg limits at Teff 5500: 3.5-5.0
    assert 5000.0 < x1[1, 0] < 6000.0 and 3.5 < x1[1, 2] < 4.5  # inside for particle 1
    np.testing.assert_array_equal(seen[1], x1)
    # costs 0.5 (particle 0: new personal and global best) and 3.0 (particle 1: no update)
    np.testing.assert_array_equal(trace[1], [0.5, *x1[0]])

    # iteration 2: particle 1's personal best is still its start, the global best is
    # particle 0's iteration-1 position: both pulls act, each with its own draw and factor
    pb = np.array([x1[0], x0[1]])
    gbest = x1[0]
    x2 = x1.copy(); v2 = v1.copy()
    for q in range(2):
        for d in range(5):
            pull = _rand_range(0.0, pb[q, d] - x2[q, d], _u(key, 2, 2, q, d, 0)) * cog
            pullg = _rand_range(0.0, gbest[d] - x2[q, d], _u(key, 2, 2, q, d, 1)) * soc
            v2[q, d] = v2[q, d] * inertia + pull + pullg
            lo, hi = {0: (5000.0, 7000.0), 1: (-1.0, 0.5), 2: (3.0, 5.0), 3: (0.0, 300.0),
                      4: (-150.0, 150.0)}[d]
            x2[q, d] = min(max(x2[q, d] + v2[q, d], lo), hi)
    # particle 0 moved on by its velocity alone: Teff 5500 + 900 at logg 3.5 (columns
    # 5000-6250 reach logg 3.5) is clamped to 6250; logg 3.5 - 0.1 at the new Teff 6250 (a
    # node: its limits (1,3) and its right neighbour's (2,3) give logg 4.0-4.5) to 4.0
    x2[0, 0] = 6250.0
    x2[0, 2] = 4.0
    np.testing.assert_allclose(seen[2], x2, rtol=0, atol=0)
    # both new costs tie the global best 0.5: with the strict "<" the global best stays
    # particle 0's iteration-1 position (particle 1 still improves its own best, 2.0 -> 0.5)
    np.testing.assert_array_equal(trace[2], [0.5, *x1[0]])
    assert bc == 0.5 and np.array_equal(best, x1[0])
    # the plausible slips give other positions: the two factors swapped, or the two draws
    for fc, fs, k1, k2 in ((soc, cog, 0, 1), (cog, soc, 1, 0)):
        alt = x1[1].copy()
        va = v1[1].copy()
        for d in (0, 3, 4):
            va[d] = (va[d] * inertia
                     + _rand_range(0.0, pb[1, d] - alt[d], _u(key, 2, 2, 1, d, k1)) * fc
                     + _rand_range(0.0, gbest[d] - alt[d], _u(key, 2, 2, 1, d, k2)) * fs)
            alt[d] = alt[d] + va[d]
        assert np.abs(alt[[0, 3, 4]] - seen[2][1, [0, 3, 4]]).max() > 1e-3


# --------------------------------------------------------------------------- build_continuum
def _map_range(x, fl, fh, tl, th):
    return (x - fl) / (fh - fl) * (th - tl) + tl


def _blend(x):
    return np.sin(np.clip(x, 0.0, 1.0) * np.pi / 2) ** 2


def _build_fp64(wl, ss, params, p1):
    """build_continuum_from_chunks in FP64: chunk polynomials on x^ in [-1, 1], then each
    chunk c >= 1 blended in with weight 1 - fac, fac = blend(map(wl; wl[stop_{c-1} - 1] ->
    0, wl[start_c] -> 1)) on the running continuum."""
    out = np.zeros(len(wl))
    for c, (a, e) in enumerate(ss):
        xh = _map_range(wl[a:e], wl[a], wl[e - 1], -1.0, 1.0)
        poly = np.polynomial.polynomial.polyval(xh, params[c * p1:(c + 1) * p1].astype(np.float64))
        if c == 0:
            out[a:e] = poly
            continue
        fac = _blend(_map_range(wl[a:e], wl[ss[c - 1][1] - 1], wl[a], 0.0, 1.0))
        out[a:e] = out[a:e] * fac + poly * (1 - fac)
    return out


@pytest.mark.parametrize("n_chunks,p_order,overlap", [(5, 4, 0.2), (10, 8, 0.2), (3, 2, 0.45)])
def test_build_continuum_fp64_and_library(n_chunks, p_order, overlap):
    from pasta_amd import _lib
    n = 6000
    wl = 10.0 ** np.linspace(np.log10(5000.0), np.log10(5500.0), n)
    rng = np.random.default_rng(n_chunks)
    p1 = p_order + 1
    params = (rng.normal(0, 1, n_chunks * p1) * 0.1).astype(np.float32)
    params[::p1] += 1e4
    oc = O.Continuum("chunk", n, wl=wl, n_chunks=n_chunks, p_order=p_order, overlap=overlap)
    ref32 = np.zeros(n, np.float32)
    O._L.or_build_continuum(oc._h, params.ctypes.data_as(C.POINTER(C.c_float)),
                            ref32.ctypes.data_as(C.POINTER(C.c_float)))
    h = C.c_void_p()
    _lib.check(_lib.lib.pasta_cont_create_chunk(wl.ctypes.data_as(C.POINTER(C.c_double)), n,
                                                n_chunks, p_order, overlap, C.byref(h)))
    try:
        ss = np.zeros(2 * n_chunks, np.int32)
        _lib.lib.pasta_cont_chunks(h, ss.ctypes.data_as(C.POINTER(C.c_int)), 2 * n_chunks)
        ss = ss.reshape(-1, 2)
        lib32 = np.zeros(n, np.float32)
        rc = _lib.lib.pasta_cont_build(h, params.ctypes.data_as(C.POINTER(C.c_float)), len(params),
                                       lib32.ctypes.data_as(C.POINTER(C.c_float)), n)
        assert rc >= 0, _lib.last_error()
    finally:
        _lib.lib.pasta_cont_destroy(h)
    exact = _build_fp64(wl, ss, params, p1)
    np.testing.assert_allclose(ref32, exact, rtol=2e-6)
    np.testing.assert_allclose(lib32, exact, rtol=2e-6)
    np.testing.assert_array_equal(lib32, ref32)
    # the blend direction matters: weighting the new chunk by fac instead fails
    wrong = exact.copy()
    for c in range(1, n_chunks):
        a, e = ss[c]
        xh = _map_range(wl[a:e], wl[a], wl[e - 1], -1.0, 1.0)
        poly = np.polynomial.polynomial.polyval(xh, params[c * p1:(c + 1) * p1].astype(np.float64))
        fac = _blend(_map_range(wl[a:e], wl[ss[c - 1][1] - 1], wl[a], 0.0, 1.0))
        prev = _build_fp64(wl, ss[:c], params, p1)[a:e] if c > 0 else 0
        wrong[a:e] = prev * (1 - fac) + poly * fac
    assert np.abs(wrong - exact).max() > 1e-3
