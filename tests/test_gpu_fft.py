"""Rotation by FFT convolution on the GPU (kernels of at least the grid's crossover taps;
the reference convolves every rotation kernel by an f32 FFT, src/convolve_rv.rs:484-529).

The path is forced on (InMemInterpolator.set_fft_taps(3): every kernel of 3 or more taps
whose tiles fit one 4096-point transform) or off (a crossover above 2,048 taps), so the same
labels run through the FFT and through the direct matrix-core paths.  Tolerance: models
within 1e-5 relative per pixel of the oracle plus a floor of 5e-7 of the spectrum's maximum
(the absolute rounding noise of an f32 FFT convolution, in the oracle's own FFT as in the
reference's and the GPU's); chi2 within 1e-5 relative; a batch bit-identical to its
evaluations alone.  Fixed resolution (instrumental kernel composed into the rotation taps),
variable resolution (a separate instrumental stage: hd > 0 pixels of halo per tile) and no
instrumental broadening.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-5
FFT_FLOOR = 5e-7
# rotation kernels of 7, 59, 279, 649, 1,105 and 1,839 taps on the small grid
VSINI = (2.0, 30.0, 150.0, 352.0, 600.0, 1000.0)
DIRECT = 1 << 20


def _dispersions(small, pasta, wlg, kind):
    wl = small.wl
    if kind == "fixed":
        return (pasta.FixedResolutionDispersion(wl, 85_000.0, wlg),
                O.Dispersion(small.og, O.DISP_FIXED, wl, resolution=85_000.0))
    if kind == "variable":
        R = np.interp(wl, [wl[0], wl[-1]], [60_000.0, 90_000.0])
        disp = (wl / R).astype(np.float32)
        return (pasta.VariableResolutionDispersion(wl, disp, wlg),
                O.Dispersion(small.og, O.DISP_VARIABLE, wl, disp=disp))
    return pasta.NoConvolutionDispersion(wl, wlg), O.Dispersion(small.og, O.DISP_NONE, wl)


def _with_taps(interp, taps, fn):
    interp.set_fft_taps(taps)
    try:
        return fn()
    finally:
        interp.set_fft_taps(0)


def _check(out, ref, what):
    ref = np.asarray(ref, np.float64)
    err = np.abs(np.asarray(out, np.float64) - ref)
    bound = RTOL * np.abs(ref) + FFT_FLOOR * np.abs(ref).max()
    assert np.all(err <= bound), (what, float(np.max(err / bound)))
    return float(np.max(err) / np.abs(ref).max())


@pytest.mark.parametrize("kind", ["fixed", "variable", "none"])
def test_fft_models_match_oracle_and_direct(small, gpu_small, kind):
    pasta, wlg, interp = gpu_small
    d, od = _dispersions(small, pasta, wlg, kind)
    labs = np.array([[6123.0, -0.3, 4.21, v, rv] for v in VSINI for rv in (-60.0, 35.0)])
    dvelo = np.log(10.0) * small.spec.log10_step
    ks = [len(O.rotation_kernel(v, dvelo)) for v in VSINI]
    print(f"fft [{kind}]: taps {ks}", flush=True)
    fft = _with_taps(interp, 3, lambda: interp.produce_model_bulk(d, labs))
    direct = _with_taps(interp, DIRECT, lambda: interp.produce_model_bulk(d, labs))
    worst = worst_d = 0.0
    for k, lab in enumerate(labs):
        ref = O.produce_model(small.og, od, lab)
        worst = max(worst, _check(fft[k], ref, ("fft", kind, lab.tolist())))
        worst_d = max(worst_d, _check(fft[k], direct[k], ("fft vs direct", kind, lab.tolist())))
    print(f"fft [{kind}]: max err / max |flux| vs oracle {worst:.2e}, vs direct {worst_d:.2e}",
          flush=True)


def test_fft_synthetic_grid_output(small, gpu_small):
    """interpolate_and_convolve (the synthetic-grid segment, no resample) on the FFT path."""
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    for v in (150.0, 600.0):
        lab = [5500.0, 0.1, 4.6, v]
        out = _with_taps(interp, 3, lambda: interp.interpolate_and_convolve(d, *lab))
        ref = O.interpolate_and_convolve(small.og, small.odisp, lab)
        assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < RTOL, lab


def test_fft_chi2_and_batch_identity(small, gpu_small):
    pasta, wlg, interp = gpu_small
    from pasta_amd import synth
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.flux,
                      small.var)
    labs = synth.random_labels(small.spec, 24, 71, vsini=(0.0, 1000.0), rv=(-100.0, 100.0))

    def run():
        chi, st, _ = star.chi2(labs)
        alone = np.array([star.chi2(labs[k:k + 1])[0][0] for k in range(len(labs))])
        return chi, st, alone

    chi, st, alone = _with_taps(interp, 3, run)
    assert not st.any()
    np.testing.assert_array_equal(chi, alone)
    ref = O.cost_batch(small.og, small.odisp, small.ocont, small.flux, small.var, labs, 8)
    rel = np.abs(chi - ref) / ref
    print(f"fft chi2: max rel {rel.max():.2e} over {len(labs)} labels", flush=True)
    assert rel.max() < RTOL


CROSSOVER = 1537  # the default: composed taps (rotation + 6 instrumental) from which FFT runs


def _vsini_for_taps(dvelo, target):
    """Smallest vsini whose rotation kernel has at least `target` taps (bisection)."""
    lo, hi = 1.0, 1500.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if len(O.rotation_kernel(mid, dvelo)) >= target:
            hi = mid
        else:
            lo = mid
    return hi


def test_fft_default_crossover(small, gpu_small):
    """With the default crossover, a kernel of 1,537 taps or more (the rotation kernel plus
    the 6 the composed R = 85,000 instrumental kernel adds; above the f16 path's 1,536) runs
    on the FFT path and a shorter one directly: each result equals the forced path's bit for
    bit."""
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    dvelo = np.log(10.0) * small.spec.log10_step
    vs = [150.0, _vsini_for_taps(dvelo, 1521), _vsini_for_taps(dvelo, 1531),
          _vsini_for_taps(dvelo, 1541), 1000.0]
    labs = np.array([[6123.0, -0.3, 4.21, v, 5.0] for v in vs])
    ks = [len(O.rotation_kernel(v, dvelo)) + 6 for v in labs[:, 3]]
    assert min(ks) < CROSSOVER <= max(ks) and sum(k >= CROSSOVER for k in ks) >= 2, ks
    default = interp.produce_model_bulk(d, labs)
    fft = _with_taps(interp, 3, lambda: interp.produce_model_bulk(d, labs))
    direct = _with_taps(interp, DIRECT, lambda: interp.produce_model_bulk(d, labs))
    for k, K in enumerate(ks):
        np.testing.assert_array_equal(default[k], fft[k] if K >= CROSSOVER else direct[k])
