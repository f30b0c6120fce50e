"""Grid loaders (src/model_fetchers.rs:16-67,136-225): file-name labels, the
includes_factor header, the directory reader and the .tar.zst reader agree on the same
files (host-side parsing; the GPU test uploads both)."""
import io
import os
import tarfile

import numpy as np
import pytest

from pasta_amd import api as A
from pasta_amd import synth


def _name(teff, m, logg):
    """filename_from_labels (src/model_fetchers.rs:39-46)"""
    mi = int(round(m * 100))
    return f"l{'m' if mi < 0 else 'p'}{abs(mi):04d}_{int(round(teff)):05d}_{int(round(logg * 100)):04d}.npy"


def _write_grid(tmp, spectra, labels, factors=None):
    d = os.path.join(tmp, "grid")
    os.makedirs(d)
    for k, (lab, row) in enumerate(zip(labels, spectra)):
        a = row
        if factors is not None:
            head = np.frombuffer(np.float32(factors[k] * 65535.0).astype("<f4").tobytes(), "<u2")
            a = np.concatenate([head, row])
        np.save(os.path.join(d, _name(*lab)), a.astype(np.uint16))
    return d


def _tar_zstd(d, path, extra=True):
    import pyarrow as pa
    buf = io.BytesIO()
    with tarfile.open(fileobj=buf, mode="w") as t:
        t.add(d, arcname="grid")
        if extra:
            info = tarfile.TarInfo("grid/README.txt")
            data = b"not a spectrum"
            info.size = len(data)
            t.addfile(info, io.BytesIO(data))
    with pa.output_stream(path, compression="zstd") as out:
        out.write(buf.getvalue())


def test_label_from_filename():
    assert A.label_from_filename("lp0020_08000_0430.npy") == (8000.0, 0.2, 4.3)
    assert A.label_from_filename("lm0150_05750_0325.npy") == (5750.0, -1.5, 3.25)
    for bad in ("lp020_08000_0430.npy", "lp0020_8000_0430.npy", "x.npy", "lp0020_08000.npy",
                "lp00a0_08000_0430.npy"):
        with pytest.raises(ValueError):
            A.label_from_filename(bad)


@pytest.fixture(scope="module")
def tiny():
    spec = synth.GridSpec(np.array([5000.0, 5250.0, 5500.0]), np.array([-0.5, 0.0]),
                          np.array([4.0, 4.5]), 512, 5000.0, 5050.0, 10)
    labels = spec.labels()
    spectra, factors = synth.make_grid(spec, includes_factor_seed=3)
    return spec, labels, spectra, factors


@pytest.mark.parametrize("with_factor", [False, True])
def test_dir_and_tar_zstd_agree(tmp_path, tiny, with_factor):
    spec, labels, spectra, factors = tiny
    d = _write_grid(str(tmp_path), spectra, labels, factors if with_factor else None)
    tz = os.path.join(str(tmp_path), "grid.tar.zst")
    _tar_zstd(d, tz)
    with open(tz, "rb") as f:
        assert f.read(4) == bytes.fromhex("28b52ffd")  # a zstd frame
    ld, sd, fd = A._read_grid_dir(d, with_factor)
    lt, st, ft = A._read_grid_tar_zstd(tz, with_factor)
    od, ot = np.lexsort(np.array(ld).T), np.lexsort(np.array(lt).T)
    assert np.array_equal(np.array(ld)[od], np.array(lt)[ot])
    assert np.array_equal(np.stack(sd)[od], np.stack(st)[ot])
    ref = {tuple(l): k for k, l in enumerate(labels)}
    for lab, row in zip(ld, sd):
        assert np.array_equal(row, spectra[ref[tuple(lab)]])
    if with_factor:
        for lab, f in zip(lt, ft):
            assert f == np.float32(np.float32(factors[ref[tuple(lab)]] * 65535.0)) / np.float32(65535.0)
    else:
        assert all(f is None for f in ft)


def test_invalid_member_name_raises(tmp_path, tiny):
    spec, labels, spectra, _ = tiny
    d = _write_grid(str(tmp_path), spectra[:2], labels[:2])
    np.save(os.path.join(d, "lbad_name.npy"), spectra[0])
    tz = os.path.join(str(tmp_path), "bad.tar.zst")
    _tar_zstd(d, tz, extra=False)
    with pytest.raises(ValueError):
        A._read_grid_tar_zstd(tz, False)


def test_negative_factor_raises():
    a = np.concatenate([np.frombuffer(np.float32(-2.0).tobytes(), "<u2"), np.zeros(4, np.uint16)])
    with pytest.raises(ValueError):
        A.read_spectrum(a, True)
