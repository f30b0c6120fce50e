"""CPU checks of the boundary: the library loads and exports every symbol the
header declares (no compute call: there is no GPU here), the ctypes binding
covers them, and host-side logic that needs no device behaves."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pasta_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pasta_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(ROOT, "pasta_amd", "libpasta_hip.so"))
    names = header_functions()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    from pasta_amd import _lib
    declared = set(header_functions())
    bound = set(_lib.SIGNATURES) | {"pasta_last_error"}
    assert declared == bound, (declared - bound, bound - declared)


def test_struct_layouts_match_header():
    from pasta_amd import _lib
    # pasta_fit_settings: 2 int + 4 double + 2x2 double + 5 int + 2x5 double
    assert C.sizeof(_lib.FitSettings) == 8 + 32 + 32 + 24 + 80
    assert C.sizeof(_lib.FitResult) == 40 + 8 + 4 + 4 + 8 + 4 + 512 * 4 + 4
    assert C.sizeof(_lib.ServerConfig) == 32


def test_last_error_without_device():
    from pasta_amd import _lib
    h = C.c_void_p()
    # argument validation happens before any HIP call
    rc = _lib.lib.pasta_cont_create_chunk(None, 0, 5, 8, 0.2, C.byref(h))
    assert rc == -1
    assert "chunk" in _lib.last_error()
    rc = _lib.lib.pasta_cont_create_constant(0, C.byref(h))
    assert rc == -1


def test_import_is_loud_without_library(tmp_path, monkeypatch):
    import importlib.util
    spec = importlib.util.spec_from_file_location("plib", os.path.join(ROOT, "pasta_amd", "_lib.py"))
    mod = importlib.util.module_from_spec(spec)
    monkeypatch.setattr(os.path, "exists", lambda p: False if p.endswith("libpasta_hip.so") else True)
    try:
        spec.loader.exec_module(mod)
        raise AssertionError("import succeeded without the library")
    except ImportError as e:
        assert "no CPU fallback" in str(e)


def test_python_api_surface():
    import pasta_amd as pasta
    for name in ["WlGrid", "InMemInterpolator", "PSOSettings", "Label", "OptimizationResult",
                 "NoConvolutionDispersion", "FixedResolutionDispersion",
                 "VariableResolutionDispersion", "ChunkContinuumFitter",
                 "LinearModelContinuumFitter", "FixedContinuum", "ConstantContinuum",
                 "FixConstraint", "RangeConstraint", "get_vsini_kernel", "InMemSingleFitter",
                 "FitServer", "LabelUncertainties"]:
        assert hasattr(pasta, name), name
    s = pasta.PSOSettings(10, 20)
    assert (s.inertia_factor, s.cognitive_factor, s.social_factor, s.delta) == (
        0.7213475204, 1.1931471806, 1.1931471806, 1e-7)
    r = pasta.OptimizationResult(pasta.Label(1, 2, 3, 4, 5), [1.0], 2.0, 3, 4.0)
    assert pasta.OptimizationResult.from_dict(r.to_dict()).to_dict() == r.to_dict()
    w = pasta.WlGrid(np.log10(5000.0), 1e-5, 10, log=True)
    np.testing.assert_allclose(w.get_array(), 10 ** (np.log10(5000.0) + 1e-5 * np.arange(10)))


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert p.returncode == 2 and "one GPU" in p.stderr
