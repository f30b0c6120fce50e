"""Grids loaded from a directory of .npy files and from a .tar.zst archive
(src/model_fetchers.rs:136-225) are resident in HBM with the same rows and factors as the
arrays they were written from."""
import ctypes as C
import os

import numpy as np
import pytest

from tests.test_loaders import _tar_zstd, _write_grid

pytestmark = pytest.mark.gpu


def _rows(interp, n_models, n_pix):
    from pasta_amd._lib import lib, ptr
    out = np.zeros((n_models, n_pix), np.uint16)
    assert lib.pasta_grid_read_rows(interp._h, 0, n_models, ptr(out, C.c_uint16)) == 0
    return out


@pytest.mark.parametrize("with_factor", [False, True])
def test_loaded_grids_match_arrays(tmp_path, small, with_factor):
    import pasta_amd as pasta
    from pasta_amd import synth
    spec = small.spec
    spectra, factors = synth.make_grid(spec, includes_factor_seed=5 if with_factor else None)
    d = _write_grid(str(tmp_path), spectra, small.labels, factors)
    tz = os.path.join(str(tmp_path), "grid.tar.zst")
    _tar_zstd(d, tz)
    wlg = pasta.WlGrid(spec.log10_first, spec.log10_step, spec.n_pix, log=True)
    ref = pasta.InMemInterpolator.from_arrays(spectra, small.labels, wlg,
                                              factors=np.float32(factors * np.float32(65535.0)) /
                                              np.float32(65535.0) if with_factor else None)
    a = pasta.InMemInterpolator(d, with_factor, wlg)
    b = pasta.InMemInterpolator.from_tar_zstd(tz, with_factor, wlg)
    n = len(small.labels)
    r0 = _rows(ref, n, spec.n_pix)
    assert np.array_equal(_rows(a, n, spec.n_pix), r0)
    assert np.array_equal(_rows(b, n, spec.n_pix), r0)
    labs = [[6123.0, -0.3, 4.21], [5400.0, 0.1, 4.6]]
    for lab in labs:
        x = ref.interpolate(*lab)
        assert np.array_equal(a.interpolate(*lab), x)
        assert np.array_equal(b.interpolate(*lab), x)
