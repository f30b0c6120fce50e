"""compute_uncertainty on the GPU (batched Brent searches over pasta_eval_chi2-style
evaluations) against the oracle's restatement (src/fitting.rs:482-568).  The two
searches see chi2 values that differ at the 1e-7 level, so their roots agree to the
Brent tolerance (1e-3 of the search radius), not bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _radius(label, radius):
    sr = np.array([2000.0, 0.3, 0.3, 40.0, 40.0] if radius is None else radius, float)
    sr[0] *= label[0]
    sr[3] = max(label[3] * sr[3], 100.0)
    return sr


@pytest.mark.parametrize("radius", [None, (0.01, 0.1, 0.1, 1.0, 1.0)])
def test_uncertainty_matches_oracle(small, gpu_small, radius):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    fitter = interp.get_fitter(d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2),
                               pasta.PSOSettings(10, 10), (0.0, 300.0), (-150.0, 150.0))
    label = small.true
    u = fitter.compute_uncertainty(interp, small.flux, small.var, 85_000.0, label, radius)
    ref, rst = O.compute_uncertainty(small.og, small.odisp, small.ocont, small.flux, small.var,
                                     85_000.0, label, radius)
    sr = _radius(label, radius)
    print(f"uncertainty {radius}: {u.to_dict()}", flush=True)
    for i, name in enumerate(pasta.LabelUncertainties.NAMES):
        got = getattr(u, name)
        for k in range(2):
            assert (got[k] is None) == bool(rst[i, k]), (name, k, got, rst[i])
            if got[k] is not None:
                assert abs(got[k] - ref[i, k]) <= 2e-3 * sr[i], (name, k, got[k], ref[i, k])


def test_uncertainty_bulk_equals_single(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    fitter = interp.get_fitter(d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2),
                               pasta.PSOSettings(10, 10), (0.0, 300.0), (-150.0, 150.0))
    labels = np.array([small.true, [6000.0, -0.2, 4.3, 15.0, 18.0], [6200.0, -0.35, 4.1, 8.0, 22.0]])
    radius = (0.01, 0.1, 0.1, 1.0, 1.0)
    bulk = fitter.compute_uncertainty_bulk(interp, [small.flux] * 3, [small.var] * 3, 85_000.0,
                                           labels, radius)
    for k in range(3):
        one = fitter.compute_uncertainty(interp, small.flux, small.var, 85_000.0, labels[k], radius)
        assert one.to_dict() == bulk[k].to_dict(), k
