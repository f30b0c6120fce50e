"""GPU parity: the HIP path against the C oracle on identical inputs (small config).

Tolerance: model spectra and chi2 within 1e-5 relative in FP32 (this path's parity
bar).  The oracle restates the reference's f32 arithmetic; the HIP
path differs in summation order (direct convolution vs the reference's FFT,
fused multiply-adds, f64 reductions), which the tolerance absorbs.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-5
FFT_FLOOR = 5e-7  # absolute rounding noise of an f32 FFT convolution, relative to its maximum


def _labels_inside(spec, n, seed, vsini=(0.0, 300.0), rv=(-150.0, 150.0)):
    from pasta_amd import synth
    return synth.random_labels(spec, n, seed, vsini=vsini, rv=rv)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


def test_interpolate_matches_oracle(small, gpu_small):
    pasta, wlg, interp = gpu_small
    rng = np.random.default_rng(11)
    spec = small.spec
    labs = [
        (5000.0, -1.0, 3.5), (7000.0, 0.5, 5.0), (6000.0, 0.0, 4.0),  # nodes and corners
        (5010.0, -0.99, 3.51), (6990.0, 0.49, 4.99),                    # quadratic edges
        (5250.0, -0.75, 4.5), (6875.3, 0.31, 3.77),
    ]
    labs += [tuple(x) for x in np.stack([rng.uniform(5000, 7000, 40), rng.uniform(-1, 0.5, 40),
                                         rng.uniform(3.5, 5.0, 40)], 1)]
    print(f"interpolate: {len(labs)} labels, W={spec.n_pix}", flush=True)
    out, st = interp.interpolate_bulk(np.array(labs))
    assert not st.any()
    for k, lab in enumerate(labs):
        ref = small.og.interpolate(*lab)
        assert _rel(out[k], ref) < RTOL, lab
        # tail quirk: the last W % 128 pixels carry the extra 1/65535
        tail = spec.n_pix - (spec.n_pix // 128) * 128
        assert np.all(out[k][-tail:] < 1e-3)


def test_interpolate_out_of_grid_errors(gpu_small):
    pasta, wlg, interp = gpu_small
    out, st = interp.interpolate_bulk(np.array([[4999.0, 0.0, 4.0], [6000.0, 0.6, 4.0],
                                                [6000.0, 0.0, 5.01]]))
    assert (st == 1).all()
    with pytest.raises(ValueError):
        interp.interpolate(4000.0, 0.0, 4.0)


@pytest.mark.parametrize("disp_kind", ["fixed", "none", "variable"])
def test_produce_model_matches_oracle(small, gpu_small, disp_kind):
    pasta, wlg, interp = gpu_small
    wl = small.wl
    if disp_kind == "fixed":
        d = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
        od = small.odisp
    elif disp_kind == "none":
        d = pasta.NoConvolutionDispersion(wl, wlg)
        od = O.Dispersion(small.og, O.DISP_NONE, wl)
    else:
        disp = (wl / np.interp(wl, [wl[0], wl[-1]], [60_000.0, 90_000.0])).astype(np.float32)
        d = pasta.VariableResolutionDispersion(wl, disp, wlg)
        od = O.Dispersion(small.og, O.DISP_VARIABLE, wl, disp=disp)
    labs = np.array([[6123.0, -0.3, 4.21, 10.0, 20.0], [6123.0, -0.3, 4.21, 0.0, 0.0],
                     [5300.0, 0.2, 3.9, 300.0, -150.0], [6800.0, -0.9, 4.9, 55.5, 149.0],
                     [5000.0, -1.0, 3.5, 1.0, -3.0]])
    labs = np.concatenate([labs, _labels_inside(small.spec, 12, 5)])
    print(f"produce_model[{disp_kind}]: {len(labs)} labels", flush=True)
    out = interp.produce_model_bulk(d, labs)
    for k, lab in enumerate(labs):
        ref = O.produce_model(small.og, od, lab)
        assert _rel(out[k], ref) < RTOL, (disp_kind, lab.tolist(), _rel(out[k], ref))


@pytest.mark.parametrize("margin", [2, 3, 4, 5, 6, 450])
def test_composed_instrumental_taps_match_oracle(small, gpu_small, margin):
    """With a fixed instrumental kernel (7 taps at R = 85,000, half-width hd = 3) the GPU
    folds it into the rotation taps when the convolved range of the evaluation stays hd
    pixels inside the grid, and convolves in two stages (zero padding between them at the
    grid edges) otherwise.  Observed ranges starting 2-6 synthetic pixels inside the grid
    put evaluations on both sides of that switch; all must match the oracle, which always
    convolves in two stages.  (At 1 pixel the last resample index reaches the end of the
    model: the reference's index panic, an error on both sides.)"""
    from pasta_amd import synth
    pasta, wlg, interp = gpu_small
    wl = synth.observed_wavelengths(small.spec, 15_000, margin)
    d = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
    od = O.Dispersion(small.og, O.DISP_FIXED, wl, resolution=85_000.0)
    rvs = (0.0, -150.0, 150.0) if margin > 200 else (0.0,)
    labs = np.array([[6123.0, -0.3, 4.21, v, rv] for v in (1.0, 10.0, 120.0, 300.0) for rv in rvs])
    print(f"composed taps: margin {margin}, {len(labs)} labels", flush=True)
    out = interp.produce_model_bulk(d, labs)
    for k, lab in enumerate(labs):
        ref = O.produce_model(small.og, od, lab).astype(np.float64)
        # Near the right edge the observed pixels reach the last W % 128 synthetic pixels,
        # which the interpolation's tail quirk scales by an extra 1/65535 (values ~1.5e-5).
        # There the oracle's f32 FFT convolution (like the reference's) carries its absolute
        # rounding noise, ~1e-7 of the spectrum's maximum, so the bound adds a floor of
        # 5e-7 of the maximum to the 1e-5 relative bar.
        err = np.abs(out[k].astype(np.float64) - ref)
        bound = RTOL * np.abs(ref) + FFT_FLOOR * np.abs(ref).max()
        assert np.all(err <= bound), (margin, lab.tolist(), float(np.max(err / bound)))


def test_interpolate_and_convolve_matches_oracle(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    for lab in ([6123.0, -0.3, 4.21, 10.0], [5500.0, 0.1, 4.6, 250.0], [6500.0, -0.5, 3.6, 0.0]):
        out = interp.interpolate_and_convolve(d, *lab)
        ref = O.interpolate_and_convolve(small.og, small.odisp, lab)
        # the segment's first/last pixels see only half a kernel: compare relative to max
        assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < RTOL, lab


def test_rotation_paths_match_oracle(small, gpu_small):
    """Rotation kernels of up to 1,536 taps (composed with the instrumental kernel) run on
    f16 matrix cores (two-way split), wider ones on f32 matrix cores (the FFT path off):
    kernels on both sides of the switch (1,521 and 1,541 rotation taps, 1,527 and 1,547
    composed) and well inside each (~626 and ~1,840 taps on this grid)."""
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    dvelo = np.log(10.0) * small.spec.log10_step
    vs = [340.0, _vsini_for_taps(dvelo, 1521), _vsini_for_taps(dvelo, 1541), 1000.0]
    labs = np.array([[6123.0, -0.3, 4.21, v, 5.0] for v in vs])
    ks = [len(O.rotation_kernel(v, dvelo)) for v in labs[:, 3]]
    assert ks[0] < 1521 and ks[1] == 1521 and ks[2] == 1541 and ks[3] > 1541, ks
    interp.set_fft_taps(1 << 20)
    print(f"rotation paths: taps {ks}", flush=True)
    try:
        out = interp.produce_model_bulk(d, labs)
    finally:
        interp.set_fft_taps(0)
    for k, lab in enumerate(labs):
        ref = O.produce_model(small.og, small.odisp, lab)
        assert _rel(out[k], ref) < RTOL, (lab.tolist(), _rel(out[k], ref))


def _vsini_for_taps(dvelo, target):
    """Smallest vsini whose rotation kernel has at least `target` taps (bisection)."""
    lo, hi = 1.0, 1500.0  # 2,757 taps at 1,500 km/s on this grid
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if len(O.rotation_kernel(mid, dvelo)) >= target:
            hi = mid
        else:
            lo = mid
    return hi


def test_rotation_kernel_limit(small, gpu_small):
    """The largest kernel the reference convolves (2,047 taps: FFTSIZE 2,048 minus one,
    made odd) matches the oracle; one of 2,049 taps is the reference's "larger than
    FFTSIZE" error (src/convolve_rv.rs:512-519), on the GPU as in the oracle."""
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    dvelo = np.log(10.0) * small.spec.log10_step
    v_max = _vsini_for_taps(dvelo, 2047)
    v_over = _vsini_for_taps(dvelo, 2049)
    assert len(O.rotation_kernel(v_max, dvelo)) == 2047
    assert len(O.rotation_kernel(v_over, dvelo)) == 2049
    print(f"rotation limit: vsini {v_max:.3f} (2047 taps), {v_over:.3f} (2049 taps)", flush=True)
    lab = np.array([[6123.0, -0.3, 4.21, v_max, -15.0]])
    out = interp.produce_model_bulk(d, lab)
    ref = O.produce_model(small.og, small.odisp, lab[0])
    assert _rel(out[0], ref) < RTOL, _rel(out[0], ref)
    _, st = interp._produce(d, np.array([[6123.0, -0.3, 4.21, v_over, -15.0]]))
    assert st[0] == 4, st  # rotation kernel larger than FFTSIZE
    with pytest.raises(O.OracleError):
        O.produce_model(small.og, small.odisp, [6123.0, -0.3, 4.21, v_over, -15.0])


def _cont_pair(small, pasta, kind):
    n = len(small.wl)
    if kind == "chunk":
        return (pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.ocont)
    if kind == "chunk10x8":
        return (pasta.ChunkContinuumFitter(small.wl, 10, 8, 0.2),
                O.Continuum("chunk", n, wl=small.wl, n_chunks=10, p_order=8, overlap=0.2))
    if kind in ("linear", "linear30"):
        from pasta_amd import synth
        dm = synth.smooth_design(n, 12 if kind == "linear" else 30, 7)
        return pasta.LinearModelContinuumFitter(dm), O.Continuum("linear", n, design=dm)
    if kind == "fixed":
        x = np.linspace(-1, 1, n)
        c = (1e4 * (1 + 0.05 * x)).astype(np.float32)
        return pasta.FixedContinuum(c, 30), O.Continuum("fixed", n, fixed=c, ignore=30)
    return pasta.ConstantContinuum(), O.Continuum("constant", n)


@pytest.mark.parametrize("kind", ["chunk", "chunk10x8", "linear", "linear30", "fixed", "constant"])
def test_chi2_matches_oracle(small, gpu_small, kind):
    pasta, wlg, interp = gpu_small
    cont, ocont = _cont_pair(small, pasta, kind)
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    labs = np.concatenate([[small.true, [6000.0, 0.0, 4.0, 30.0, -10.0]],
                           _labels_inside(small.spec, 30, 9)])
    print(f"chi2[{kind}]: {len(labs)} labels", flush=True)
    star = pasta.Star(interp, d, cont, small.flux, small.var)
    chi2, st, params = star.chi2(labs)
    assert not st.any()
    for k, lab in enumerate(labs):
        ref, rp = O.cost(small.og, small.odisp, ocont, small.flux, small.var, lab)
        if kind == "constant":
            # The mean of f/m is the reference's sequential f32 fold (continuum_fitting.rs:
            # 490-492) on both sides; it differs from the exactly summed mean by ~1e-5 here,
            # which moves chi2 by ~1e-4, so this pins the fold order itself.
            assert abs(params[k][0] - rp[0]) <= 2e-6 * abs(rp[0]), (params[k][0], rp[0])
        assert abs(chi2[k] - ref) / ref < RTOL, (kind, lab.tolist(), chi2[k], ref)
        if kind in ("chunk", "chunk10x8", "linear", "linear30", "constant"):
            np.testing.assert_allclose(params[k][: len(rp)], rp, rtol=2e-4, atol=1e-3 * np.abs(rp).max())


def test_linear_continuum_ragged(small, gpu_small):
    """Linear continuum (continuum_fitting.rs:69-95) with 30 columns over 12,345 pixels
    (not a whole number of pixel quads) and 70 labels (a full group of 64 evaluations and
    a partial one), including one label outside the grid."""
    from pasta_amd import synth
    pasta, wlg, interp = gpu_small
    wl = synth.observed_wavelengths(small.spec, 12_345, 500)
    od = O.Dispersion(small.og, O.DISP_NONE, wl)
    flux, var = synth.observe(O.produce_model(small.og, od, small.true), wl, 120.0, 41)
    dm = synth.smooth_design(len(wl), 30, 11)
    star = pasta.Star(interp, pasta.NoConvolutionDispersion(wl, wlg), pasta.LinearModelContinuumFitter(dm),
                      flux, var)
    labs = np.concatenate([[small.true], _labels_inside(small.spec, 68, 12),
                           [[3000.0, 0.0, 4.0, 10.0, 0.0]]])
    chi2, st, params = star.chi2(labs)
    assert st[-1] != 0 and not st[:-1].any()
    oc = O.Continuum("linear", len(wl), design=dm)
    for k, lab in enumerate(labs[:-1]):
        ref, rp = O.cost(small.og, od, oc, flux, var, lab)
        assert abs(chi2[k] - ref) / ref < RTOL, (k, lab.tolist(), chi2[k], ref)
        np.testing.assert_allclose(params[k][: len(rp)], rp, rtol=2e-4, atol=1e-3 * np.abs(rp).max())


def test_chi2_errors_are_nan(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    fitter = interp.get_fitter(d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2),
                               pasta.PSOSettings(4, 1), (0, 300), (-150, 150))
    bad = [[4000.0, 0.0, 4.0, 10.0, 0.0], [6000.0, 0.0, 4.0, -1.0, 0.0],
           [6000.0, 0.0, 4.0, 10.0, 5000.0], small.true.tolist()]
    out = fitter.chi2(interp, small.flux, small.var, bad)
    assert np.isnan(out[:3]).all() and np.isfinite(out[3])
    with pytest.raises(ValueError):
        fitter.chi2(interp, small.flux, small.var, bad, allow_nan=False)


def test_results_are_deterministic(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.flux,
                      small.var)
    labs = _labels_inside(small.spec, 24, 3)
    a, _, _ = star.chi2(labs)
    b, _, _ = star.chi2(labs)
    c, _, _ = star.chi2(labs[::-1].copy())
    assert np.array_equal(a, b)
    assert np.array_equal(a, c[::-1])


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7])
def test_same_cell_groups_match_alone_and_oracle(small, gpu_small, n):
    """Evaluations whose labels share one grid cell gather the same 64 neighbour rows, and the
    model kernel interpolates up to three of them from one gather (groups of 3, 3, ..., then 2
    or 1).  Each one's chi2 must equal, bit for bit, its chi2 evaluated alone, and match the
    oracle."""
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.flux,
                      small.var)
    rng = np.random.default_rng(40 + n)
    # inside the cell Teff (6000, 6250), [M/H] (-0.5, -0.25), logg (4.0, 4.5): same rows
    labs = np.stack([rng.uniform(6010, 6240, n), rng.uniform(-0.49, -0.26, n),
                     rng.uniform(4.01, 4.49, n), rng.uniform(0, 300, n),
                     rng.uniform(-150, 150, n)], axis=1)
    idx = [small.og.local_weights(*lab[:3])[0] for lab in labs]
    assert all(np.array_equal(idx[0], r) for r in idx), "labels must share their neighbour rows"
    print(f"same-cell groups: {n} evaluations", flush=True)
    chi2, st, _ = star.chi2(labs)
    assert not st.any()
    for k, lab in enumerate(labs):
        alone, st1, _ = star.chi2(lab[None, :].copy())
        assert not st1.any()
        assert chi2[k] == alone[0], (n, k, chi2[k], alone[0])
        ref, _ = O.cost(small.og, small.odisp, small.ocont, small.flux, small.var, lab)
        assert abs(chi2[k] - ref) / ref < RTOL, (n, lab.tolist(), chi2[k], ref)


def test_chi2_bulk_matches_oracle(small, gpu_small):
    """chi2_bulk (src/lib.rs:1297-1338): per-spectrum wavelengths, NoConvolution and
    ChunkFitter(5, 8, 0.2) hard-coded, one label each, in one batched launch."""
    from pasta_amd import synth
    pasta, wlg, interp = gpu_small
    wls = [small.wl, synth.observed_wavelengths(small.spec, 9_000, 450),
           synth.observed_wavelengths(small.spec, 12_345, 500)]
    labs = np.array([[6123.0, -0.3, 4.21, 10.0, 20.0], [5600.0, 0.1, 4.5, 80.0, -60.0],
                     [6800.0, -0.8, 3.7, 150.0, 100.0]])
    obs = []
    for k, wl in enumerate(wls):
        od = O.Dispersion(small.og, O.DISP_NONE, wl)
        obs.append(synth.observe(O.produce_model(small.og, od, labs[k]), wl, 100.0, 300 + k))
    got = interp.chi2_bulk(wls, [o[0] for o in obs], [o[1] for o in obs], labs)
    for k, wl in enumerate(wls):
        od = O.Dispersion(small.og, O.DISP_NONE, wl)
        oc = O.Continuum("chunk", len(wl), wl=wl, n_chunks=5, p_order=8, overlap=0.2)
        ref, _ = O.cost(small.og, od, oc, obs[k][0], obs[k][1], labs[k])
        assert abs(got[k] - ref) / ref < RTOL, (k, got[k], ref)
    # a failing label: NaN with allow_nan, else an error
    bad = labs.copy()
    bad[1, 0] = 20000.0
    out = interp.chi2_bulk(wls, [o[0] for o in obs], [o[1] for o in obs], bad, allow_nan=True)
    assert np.isnan(out[1]) and not np.isnan(out[0])
    with pytest.raises(Exception):
        interp.chi2_bulk(wls, [o[0] for o in obs], [o[1] for o in obs], bad)
