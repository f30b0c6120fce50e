"""The allocation-free chi2 entry point (pasta_eval_chi2_noalloc) against pasta_eval_chi2,
called through the C ABI: bit-identical chi2, statuses and continuum parameters on two
calls in a row, the second with a smaller star and fewer evaluations, for every
continuum kind; a failed evaluation reports the same status and error; an evaluation
set larger than the scratch array is rejected."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _call(fn, star, labels):
    from pasta_amd import _lib
    lab = np.ascontiguousarray(labels, np.float64)
    P = len(lab)
    chi2 = np.zeros(P)
    st = np.full(P, -7, np.int32)
    params = np.zeros((P, max(1, star.n_params)), np.float32)
    rc = fn(star._h, lab.ctypes.data_as(C.POINTER(C.c_double)), P,
            chi2.ctypes.data_as(C.POINTER(C.c_double)), st.ctypes.data_as(C.POINTER(C.c_int)),
            params.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, _lib.last_error() if rc else "", chi2, st, params


def _stars(small, pasta, wlg, interp, n, kind):
    from pasta_amd import synth
    wl = small.wl if n == len(small.wl) else synth.observed_wavelengths(small.spec, n, 450)
    d = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
    od = O.Dispersion(small.og, O.DISP_FIXED, wl, resolution=85_000.0)
    flux, var = synth.observe(O.produce_model(small.og, od, small.true), wl, 100.0, 2000 + n)
    if kind == "chunk":
        c = pasta.ChunkContinuumFitter(wl, 5, 4, 0.2)
    elif kind == "linear":
        c = pasta.LinearModelContinuumFitter(synth.smooth_design(n, 12, 3))
    elif kind == "fixed":
        c = pasta.FixedContinuum(np.full(n, 1e4, np.float32))
    else:
        c = pasta.ConstantContinuum()
    return pasta.Star(interp, d, c, flux, var)


@pytest.mark.parametrize("kind", ["chunk", "linear", "fixed", "constant"])
def test_noalloc_equals_eval_chi2_twice(small, gpu_small, kind):
    from pasta_amd import _lib, synth
    pasta, wlg, interp = gpu_small
    for call, (n, P, seed) in enumerate([(15_000, 64, 1), (9_000, 13, 2)]):
        print(f"noalloc[{kind}] call {call}: N={n} P={P}", flush=True)
        star = _stars(small, pasta, wlg, interp, n, kind)
        labels = synth.random_labels(small.spec, P, seed)
        ref = _call(_lib.lib.pasta_eval_chi2, star, labels)
        got = _call(_lib.lib.pasta_eval_chi2_noalloc, star, labels)
        assert ref[0] == 0 and got[0] == 0, (ref[1], got[1])
        np.testing.assert_array_equal(got[2], ref[2])
        np.testing.assert_array_equal(got[3], ref[3])
        np.testing.assert_array_equal(got[4], ref[4])
        assert np.all(np.isfinite(got[2]))


def test_noalloc_failed_evaluation_and_limit(small, gpu_small):
    from pasta_amd import _lib, synth
    pasta, wlg, interp = gpu_small
    star = _stars(small, pasta, wlg, interp, 15_000, "chunk")
    labels = synth.random_labels(small.spec, 5, 9)
    labels[3, 0] = 4000.0  # outside the grid
    ref = _call(_lib.lib.pasta_eval_chi2, star, labels)
    got = _call(_lib.lib.pasta_eval_chi2_noalloc, star, labels)
    assert ref[0] == got[0] == _lib.PASTA_ERR_MODEL and ref[1] == got[1]
    np.testing.assert_array_equal(got[3], ref[3])
    assert got[3][3] == 1 and np.isnan(got[2][3])
    np.testing.assert_array_equal(got[2][[0, 1, 2, 4]], ref[2][[0, 1, 2, 4]])
    # without a status array the error is still the failed evaluation's
    lab = np.ascontiguousarray(labels)
    chi2 = np.zeros(5)
    rc = _lib.lib.pasta_eval_chi2_noalloc(star._h, lab.ctypes.data_as(C.POINTER(C.c_double)), 5,
                                          chi2.ctypes.data_as(C.POINTER(C.c_double)), None, None)
    assert rc == _lib.PASTA_ERR_MODEL and _lib.last_error() == ref[1]
    # 1,200 evaluations of N = 15,000 need ~72 MB: more than the scratch array
    big = np.repeat(labels[:1], 1200, axis=0)
    rc, err, *_ = _call(_lib.lib.pasta_eval_chi2_noalloc, star, big)
    assert rc == -1 and "scratch" in err
