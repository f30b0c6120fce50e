"""Writes tests/golden/small_config.npz: oracle outputs for the small configuration.

The reference publishes no golden vectors and cannot be built here, so these fixtures
are the C oracle's own outputs on the seeded small configuration (9x7x4 grid of
20,000 px, 15,000 observed px, 5 chunks of degree 4).  They pin the oracle and the
synthetic generators against drift, and give the HIP path fixed vectors to match.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pasta_amd import synth  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "small_config.npz")


def grid_digest(spectra):
    return hashlib.sha256(np.ascontiguousarray(spectra).tobytes()).hexdigest()


def build():
    spec = synth.small_grid_spec()
    spectra, _ = synth.make_grid(spec)
    og = O.Grid(spec.labels(), spectra, None, spec.log10_first, spec.log10_step)
    wl = synth.observed_wavelengths(spec, 15_000, 450)
    od = O.Dispersion(og, O.DISP_FIXED, wl, resolution=85_000.0)
    true = np.array([6123.0, -0.3, 4.21, 10.0, 20.0])
    flux, var = synth.observe(O.produce_model(og, od, true), wl, 100.0, 2000)
    oc = O.Continuum("chunk", len(wl), wl=wl, n_chunks=5, p_order=4, overlap=0.2)
    labels = np.array([[6123.0, -0.3, 4.21, 10.0, 20.0],   # the truth
                       [5000.0, -1.0, 3.5, 0.0, 0.0],      # grid corner, no broadening
                       [6990.0, 0.49, 4.99, 300.0, -150.0],  # quadratic edges, widest kernel
                       [5620.0, -0.13, 4.37, 77.7, 101.0],
                       [6400.0, 0.2, 3.8, 150.0, -42.0]])
    models = np.stack([O.produce_model(og, od, l) for l in labels])
    chi2 = np.array([O.cost(og, od, oc, flux, var, l)[0] for l in labels])
    interp = np.stack([og.interpolate(*l[:3]) for l in labels[:3]])
    return dict(grid_sha256=np.array(grid_digest(spectra)), labels=labels, wl=wl, flux=flux,
                var=var, models=models.astype(np.float32), chi2=chi2,
                interp=interp.astype(np.float32))


if __name__ == "__main__":
    d = build()
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB)")
