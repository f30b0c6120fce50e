"""Builds libpasta_hip.so (gfx950) in-tree with hipcc.

    python pasta_amd/build.py            # libraries
    python pasta_amd/build.py --all      # + the CPU oracle (test infrastructure)

Flags: -ffp-contract=off keeps every f32/f64 expression rounding as written
(only explicit fmaf/fma fuse); packed FP32 instructions are disabled for the
device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpasta_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "api.hip", "binary.hip", "noalloc.hip"]
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build_lib(verbose=True, defines=None, out=LIB):
    """defines: {name: value} compile-time shape parameters (PASTA_TILE, PASTA_CONT_BLOCK,
    PASTA_CHI2_GROUP; csrc/engine.h) for a library built to `out` instead of the package's."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    dflags = [f"-D{k}={v}" for k, v in sorted((defines or {}).items())]
    cmd = [HIPCC] + FLAGS + dflags + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


SYNTH_LIB = os.path.join(HERE, "libpasta_synth.so")


def build_synth(verbose=True):
    tmp = SYNTH_LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, os.path.join(CSRC, "synth.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_oracle(verbose=True):
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "oracle")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def main(argv):
    build_lib()
    build_synth()
    if "--all" in argv:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv[1:])
