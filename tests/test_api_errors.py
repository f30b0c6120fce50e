"""CPU checks of the Python layer's error mapping for batched calls: a call that fails as a
whole (HIP error, allocation failure) raises instead of returning the zero-filled output
arrays, a bad argument raises ValueError, and PASTA_ERR_MODEL (per-item failures, reported
in the status array) passes through.  The library is replaced by a stub, so no GPU is needed."""
import numpy as np
import pytest

import pasta_amd as pasta
from pasta_amd import _lib, api


class _StubLib:
    def __init__(self, rc):
        self.rc = rc

    def __getattr__(self, name):
        rc = self.rc
        return lambda *a: rc


def _star():
    s = object.__new__(api.Star)
    s._h, s._own, s.n, s.n_params = None, False, 10, 3
    return s


@pytest.fixture
def stub(monkeypatch):
    def use(rc, msg):
        monkeypatch.setattr(api, "lib", _StubLib(rc))
        monkeypatch.setattr(_lib, "last_error", lambda: msg)
    return use


def test_chi2_raises_on_hip_error(stub):
    stub(_lib.PASTA_ERR_HIP, "hipMemcpyAsync: invalid argument")
    with pytest.raises(_lib.PastaError, match="invalid argument"):
        _star().chi2(np.zeros((2, 5)))


def test_chi2_raises_on_allocation_failure(stub):
    stub(_lib.PASTA_ERR_NOMEM, "hipMalloc: out of memory")
    with pytest.raises(_lib.PastaError, match="out of memory"):
        _star().chi2(np.zeros((4, 5)))


def test_chi2_bad_argument_is_value_error(stub):
    stub(_lib.PASTA_ERR_ARG, "bad arguments")
    with pytest.raises(ValueError):
        _star().chi2(np.zeros((1, 5)))


def test_chi2_model_errors_pass_through(stub):
    stub(_lib.PASTA_ERR_MODEL, "labels out of the grid bounds")
    chi2, st, params = _star().chi2(np.zeros((3, 5)))
    assert chi2.shape == (3,) and st.shape == (3,) and params.shape == (3, 3)


def test_interpolate_bulk_raises_on_hip_error(stub):
    stub(_lib.PASTA_ERR_HIP, "kernel launch failed")
    interp = object.__new__(pasta.InMemInterpolator)
    interp._h, interp.n_pix = None, 16
    with pytest.raises(_lib.PastaError, match="kernel launch failed"):
        interp.interpolate_bulk(np.zeros((2, 3)))
    with pytest.raises(_lib.PastaError):
        interp.interpolate_linear_bulk(np.zeros((2, 3)))
