"""Shape tuner (pasta_amd/tune.py): candidate table and build wiring, without a GPU or a
compiler run."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _tune():
    spec = importlib.util.spec_from_file_location("pasta_tune", os.path.join(ROOT, "pasta_amd", "tune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_candidates_respect_the_shape_rules():
    """Tiles are 256 k - 16 pixels with 3 <= k <= 15, blocks whole 1024-pixel units, chi2
    groups 4, 8 or 16 (the static_asserts of csrc/engine.h), one parameter per candidate."""
    t = _tune()
    assert list(t.CANDIDATES)[0] == "default" and t.CANDIDATES["default"] == {}
    for name, d in t.CANDIDATES.items():
        assert len(d) <= 1, name
        for k, v in d.items():
            if k == "PASTA_TILE":
                assert v % 256 == 240 and 752 <= v <= 3824, name
            elif k == "PASTA_CONT_BLOCK":
                assert v % 1024 == 0 and v >= 1024, name
            elif k == "PASTA_CHI2_GROUP":
                assert v in (4, 8, 16), name
            else:
                pytest.fail(f"unknown shape parameter {k}")
    lo = [b[0] for b in t.BUCKETS]
    hi = [b[1] for b in t.BUCKETS]
    assert lo == sorted(lo) and all(h > l for l, h in t.BUCKETS)
    assert lo[0] == 40_000 and hi[-1] == 200_000


def test_engine_defaults_are_the_default_candidate():
    """The library's built-in shapes are the ones the tuner calls default."""
    src = open(os.path.join(ROOT, "pasta_amd", "csrc", "engine.h")).read()
    t = _tune()
    for k in ("PASTA_TILE", "PASTA_CONT_BLOCK", "PASTA_CHI2_GROUP"):
        assert f"#define {k} " in src
        others = {n for n, d in t.CANDIDATES.items() if k in d}
        default = int(src.split(f"#define {k} ")[1].split()[0])
        assert all(t.CANDIDATES[n][k] != default for n in others), k


def test_build_stages_each_candidate(tmp_path, monkeypatch):
    """build() copies the package's Python files and bench.py beside one library per
    candidate, compiled with that candidate's defines (compiler call replaced)."""
    t = _tune()
    sys.path.insert(0, os.path.join(ROOT, "pasta_amd"))
    import build as B

    calls = []

    def fake_build_lib(verbose=True, defines=None, out=None):
        calls.append((dict(defines or {}), out))
        open(out, "wb").close()
        return out

    monkeypatch.setattr(B, "build_lib", fake_build_lib)
    if not os.path.exists(B.SYNTH_LIB):
        pytest.skip("libpasta_synth.so is not built")
    t.build(str(tmp_path), ["default", "tile2032"])
    assert [c[0] for c in calls] == [{}, {"PASTA_TILE": 2032}]
    for name in ("default", "tile2032"):
        d = tmp_path / name
        assert (d / "bench.py").exists()
        for f in t.PY_FILES + ["libpasta_synth.so", "libpasta_hip.so"]:
            assert (d / "pasta_amd" / f).exists(), (name, f)
