"""Work items over the union of present rows (k_union): evaluations whose stencils overlap
(the same cell, or cells one node apart along [M/H], logg or Teff) share one gather of their
rows, every evaluation multiplying each union row by its own weight (0 off its stencil), and
evaluations at the grid's edges and corners (absent neighbours: 48, 36 or 27 present rows)
gather only their present rows.  Each result must equal the same evaluation alone, bit for
bit, and the oracle within 1e-5 (src/interpolate.rs:930-983 for the interpolation,
src/fitting.rs:313-326 for the cost)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-5

# small grid: Teff 5000-7000/250, [M/H] -1..0.5/0.25, logg 3.5-5/0.5
BASE = (6123.0, -0.12, 4.21)
LABELS = [
    BASE, (6180.0, -0.20, 4.30), (6050.0, -0.05, 4.10),       # one cell
    (6123.0, 0.10, 4.21), (6123.0, -0.40, 4.21),             # [M/H] neighbours
    (6123.0, -0.12, 4.60), (6123.0, -0.12, 3.80),            # logg neighbours
    (6300.0, -0.12, 4.21), (5900.0, -0.12, 4.21),            # Teff neighbours
    (5000.0, -0.12, 4.21), (6123.0, -1.00, 4.21),            # on an edge node: 48 rows
    (6123.0, -0.12, 5.00), (5000.0, -1.00, 4.21),            # 48, 36 rows
    (5000.0, -1.00, 3.50), (7000.0, 0.50, 5.00),             # corners: 27 rows
    (6000.0, 0.00, 4.00), (6000.0, 0.00, 4.00),              # interior nodes, twice
]


def _labels():
    rng = np.random.default_rng(11)
    out = []
    for t, m, g in LABELS:
        for _ in range(2):  # two particles per position: pairs of equal stencils too
            out.append([t, m, g, rng.uniform(0.0, 120.0), rng.uniform(-60.0, 60.0)])
    return np.array(out)


def test_stencil_counts_cover_edges_and_corners(small):
    present = [int((small.og.local_weights(*lab)[0] >= 0).sum()) for lab in LABELS]
    assert {64, 48, 36, 27} <= set(present), present


def test_union_work_items_equal_alone_and_oracle(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2), small.flux,
                      small.var)
    labs = _labels()
    chi, st, _ = star.chi2(labs)
    assert not st.any()
    alone = np.array([star.chi2(labs[k:k + 1])[0][0] for k in range(len(labs))])
    np.testing.assert_array_equal(chi, alone)
    ref = O.cost_batch(small.og, small.odisp, small.ocont, small.flux, small.var, labs, 8)
    rel = np.abs(chi - ref) / ref
    print(f"union items: chi2 max rel {rel.max():.2e} over {len(labs)} labels", flush=True)
    assert rel.max() < RTOL


def test_union_work_items_models_equal_alone(small, gpu_small):
    pasta, wlg, interp = gpu_small
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    labs = _labels()
    bulk = interp.produce_model_bulk(d, labs)
    for k, lab in enumerate(labs):
        np.testing.assert_array_equal(bulk[k], interp.produce_model_bulk(d, lab[None, :])[0])
        ref = O.produce_model(small.og, small.odisp, lab)
        assert np.max(np.abs(bulk[k] - ref) / np.abs(ref)) < RTOL, lab.tolist()
