"""Small-configuration fixtures (tests/golden/small_config.npz, written by
tests/golden/make_fixtures.py from the C oracle).  The reference ships no golden
vectors and cannot be built here; these pin the oracle and the seeded generators
against drift (CPU) and give the HIP path fixed vectors to match (GPU, 1e-5)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "small_config.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_generators_unchanged(small, golden):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_fixtures import grid_digest
    assert grid_digest(small.spectra) == str(golden["grid_sha256"])
    np.testing.assert_array_equal(small.wl, golden["wl"])
    np.testing.assert_array_equal(small.flux, golden["flux"])
    np.testing.assert_array_equal(small.var, golden["var"])


def test_oracle_reproduces_fixtures(small, golden):
    from oracle import oracle as O
    for k, lab in enumerate(golden["labels"]):
        np.testing.assert_array_equal(O.produce_model(small.og, small.odisp, lab), golden["models"][k])
        c, _ = O.cost(small.og, small.odisp, small.ocont, small.flux, small.var, lab)
        assert c == golden["chi2"][k]
    for k, lab in enumerate(golden["labels"][:3]):
        np.testing.assert_array_equal(small.og.interpolate(*lab[:3]), golden["interp"][k])


@pytest.mark.gpu
def test_hip_matches_fixtures(small, gpu_small, golden):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    out = interp.produce_model_bulk(d, golden["labels"])
    for k in range(len(golden["labels"])):
        ref = golden["models"][k].astype(np.float64)
        assert np.max(np.abs(out[k] - ref) / np.maximum(np.abs(ref), 1e-30)) < 1e-5, k
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.flux, small.var)
    chi, st, _ = star.chi2(golden["labels"])
    assert not st.any()
    np.testing.assert_allclose(chi, golden["chi2"], rtol=1e-5)
    got, _ = interp.interpolate_bulk(golden["labels"][:3, :3])
    np.testing.assert_allclose(got, golden["interp"], rtol=1e-5)
