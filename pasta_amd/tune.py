"""Shape tuner for the batched PSO iteration.

The kernels' tile and block shapes are compile-time parameters (csrc/engine.h):

* PASTA_TILE        synthetic pixels per k_model workgroup (256 k - 16, k = 3..15);
* PASTA_CONT_BLOCK  observed pixels per continuum reduction block;
* PASTA_CHI2_GROUP  evaluations of one star per chi2 workgroup (4, 8 or 16).

`build` compiles one library per candidate and stages it beside a copy of the package's
Python files and of bench.py (any machine with hipcc).  `run` (on the GPU) first checks
that every candidate gives the default library's chi2 on a fixed set of evaluations of a
small-grid star, then times the batched PSO iteration of bench.py with each candidate for
every spectrum-length bucket, and prints one table row per (bucket, candidate) and the
fastest candidate of each bucket.

    python pasta_amd/tune.py build --dir DIR
    python pasta_amd/tune.py run --dir DIR [--stars 256] [--steps 20] [--out FILE]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CANDIDATES = {
    "default": {},
    "tile3568": {"PASTA_TILE": 3568},
    "tile3056": {"PASTA_TILE": 3056},
    "tile2800": {"PASTA_TILE": 2800},
    "tile2032": {"PASTA_TILE": 2032},
    "block2048": {"PASTA_CONT_BLOCK": 2048},
    "block8192": {"PASTA_CONT_BLOCK": 8192},
    "group4": {"PASTA_CHI2_GROUP": 4},
    "group16": {"PASTA_CHI2_GROUP": 16},
}
# observed pixels per star of the batched workload, U{lo..hi}
BUCKETS = [(40_000, 80_000), (80_000, 140_000), (140_000, 200_000)]
PY_FILES = ["__init__.py", "_lib.py", "api.py", "synth.py"]
CHECK_RTOL = 2e-6  # summation order only (f64 block partials, rotation tile alignment)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build(dirname, names):
    sys.path.insert(0, HERE)
    import build as B  # pasta_amd/build.py

    for name in names:
        d = os.path.join(dirname, name)
        pkg = os.path.join(d, "pasta_amd")
        os.makedirs(pkg, exist_ok=True)
        for f in PY_FILES:
            shutil.copy2(os.path.join(HERE, f), pkg)
        shutil.copy2(B.SYNTH_LIB, pkg)
        shutil.copy2(os.path.join(ROOT, "bench.py"), d)
        B.build_lib(verbose=False, defines=CANDIDATES[name],
                    out=os.path.join(pkg, "libpasta_hip.so"))
        log(f"[tune] built {name} {CANDIDATES[name]}")


def check_one(pkgdir, out):
    """chi2 of 64 fixed label vectors (vsini 0..300, rv -150..150) of one small-grid star
    with the library staged in pkgdir, saved to out (.npy)."""
    sys.path.insert(0, pkgdir)
    import numpy as np

    import pasta_amd as pasta
    from pasta_amd import synth

    spec = synth.small_grid_spec()
    spectra, _ = synth.make_grid(spec)
    wlg = pasta.WlGrid(spec.log10_first, spec.log10_step, spec.n_pix, log=True)
    interp = pasta.InMemInterpolator.from_arrays(spectra, spec.labels(), wlg)
    wl = synth.observed_wavelengths(spec, 15_000, 450)
    disp = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
    model = interp.produce_model(disp, 6123.0, -0.3, 4.21, 10.0, 20.0)
    flux, var = synth.observe(model, wl, 100.0, 2000)
    rng = np.random.default_rng(7)
    lab = np.stack([rng.uniform(5000, 7000, 64), rng.uniform(-1.0, 0.5, 64),
                    rng.uniform(3.5, 5.0, 64), rng.uniform(0, 300, 64),
                    rng.uniform(-150, 150, 64)], axis=1)
    res = []
    for cont in (pasta.ChunkContinuumFitter(wl, 5, 4, 0.2), pasta.ConstantContinuum()):
        chi2, st, _ = pasta.Star(interp, disp, cont, flux, var).chi2(lab)
        res.append(np.where(st == 0, chi2, np.nan))
    np.save(out, np.concatenate(res))


def run(dirname, names, stars, steps, out):
    import numpy as np

    ref = None
    for name in names:
        d = os.path.join(dirname, name)
        f = os.path.join(d, "check.npy")
        subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "check-one",
                        "--pkg", d, "--out", f], check=True, timeout=600)
        v = np.load(f)
        if ref is None:
            ref = v
            log(f"[tune] check {name}: reference ({np.isfinite(v).sum()} finite chi2)")
            continue
        same_nan = np.array_equal(np.isnan(v), np.isnan(ref))
        rel = np.nanmax(np.abs(v - ref) / np.abs(ref)) if same_nan else np.inf
        log(f"[tune] check {name}: max rel chi2 difference {rel:.2e}")
        if not rel <= CHECK_RTOL:
            raise SystemExit(f"candidate {name} changes chi2 by {rel:.2e} (> {CHECK_RTOL})")
    rows = []
    for lo, hi in BUCKETS:
        for name in names:
            d = os.path.join(dirname, name)
            cmd = [sys.executable, "-u", os.path.join(d, "bench.py"), "--stars", str(stars),
                   "--steps", str(steps), "--warmup", "1", "--no-plain", "--no-cpu-baseline",
                   "--no-serving", "--no-binary", "--n-range", str(lo), str(hi)]
            t0 = time.time()
            p = subprocess.run(cmd, check=True, timeout=600, stdout=subprocess.PIPE, text=True)
            r = json.loads(p.stdout.strip().splitlines()[-1])
            rows.append(dict(bucket=f"{lo}-{hi}", candidate=name, ms_per_step=r["ms_per_step"],
                             k_model_ms=r["roofline"]["kernel_ms"],
                             iteration_ms=r["roofline"]["iteration_ms"]))
            log(f"[tune] {lo}-{hi} {name}: {r['ms_per_step']:.2f} ms/step, k_model "
                f"{r['roofline']['kernel_ms']:.2f} ms ({time.time() - t0:.0f} s)")
    lines = ["| bucket (observed px) | candidate | ms per iteration | k_model ms | eager iteration ms |",
             "|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| {r['bucket']} | {r['candidate']} | {r['ms_per_step']:.2f} | "
                     f"{r['k_model_ms']:.2f} | {r['iteration_ms']:.2f} |")
    lines.append("")
    for lo, hi in BUCKETS:
        b = [r for r in rows if r["bucket"] == f"{lo}-{hi}"]
        best = min(b, key=lambda r: r["ms_per_step"])
        dflt = next(r for r in b if r["candidate"] == names[0])
        lines.append(f"best for {lo}-{hi}: {best['candidate']} ({best['ms_per_step']:.2f} ms, "
                     f"{names[0]} {dflt['ms_per_step']:.2f} ms)")
    text = "\n".join(lines)
    print(text, flush=True)
    if out:
        with open(out, "w") as fh:
            fh.write(text + "\n")


def main(argv):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("what", choices=["build", "run", "check-one"])
    ap.add_argument("--dir", default=None)
    ap.add_argument("--only", default=None, help="comma-separated candidate names")
    ap.add_argument("--stars", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--out", default=None)
    ap.add_argument("--pkg", default=None)
    a = ap.parse_args(argv)
    names = list(CANDIDATES) if not a.only else a.only.split(",")
    if names[0] != "default":
        names = ["default"] + [n for n in names if n != "default"]
    if a.what == "build":
        build(a.dir, names)
    elif a.what == "run":
        run(a.dir, names, a.stars, a.steps, a.out)
    else:
        check_one(a.pkg, a.out)


if __name__ == "__main__":
    main(sys.argv[1:])
