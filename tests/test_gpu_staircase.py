"""GPU parity on a staircase grid (per-Teff logg limits, the shape of real PASTA grids,
src/interpolate.rs:255-331,378-533) and for single-star Fix/Range constraints
(src/bounds.rs:45-60,106-130).

Grid: the small configuration's Teff (5000-7000 / 250), [M/H] (-1..0.5 / 0.25) and line
recipe (20,000 px over 5000-5500 A, 200 lines), logg 3.0-5.0 / 0.5 with index limits
(0,4) for Teff <= 5500, (1,4) for 5750-6250 and (2,4) from 6500 K: the Teff/logg
neighbour matrix loses entries at the steps, the per-row Teff and per-column logg
factors fall back to quadratic there (src/cubic.rs:159-243), and the 64 weights need
not sum to 1.

Tolerance: models and chi2 within 1e-5 relative (FP32); neighbour rows and f32 weights
identical to the oracle's; PSO traces on this grid (staircase clamping,
src/interpolate.rs:625-664) bit-identical to the oracle's PSO fed the GPU's chi2.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-5
LIMITS = [(0, 4)] * 3 + [(1, 4)] * 3 + [(2, 4)] * 3


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


@pytest.fixture(scope="module")
def stair():
    import pasta_amd as pasta
    from pasta_amd import synth
    spec = synth.small_grid_spec()
    spec.logg = np.arange(3.0, 5.01, 0.5)
    labels = []
    for i, t in enumerate(spec.teff):
        lo, hi = LIMITS[i]
        for k in range(lo, hi + 1):
            for mh in spec.m:
                labels.append((t, mh, spec.logg[k]))
    labels = np.array(labels)
    spectra, _ = synth.make_grid(spec, labels)
    og = O.Grid(labels, spectra, None, spec.log10_first, spec.log10_step)
    wlg = pasta.WlGrid(spec.log10_first, spec.log10_step, spec.n_pix, log=True)
    # rows handed over in another order than storage order: the library sorts them
    perm = np.random.default_rng(5).permutation(len(labels))
    interp = pasta.InMemInterpolator.from_arrays(spectra[perm], labels[perm], wlg)
    wl = synth.observed_wavelengths(spec, 15_000, 450)
    od = O.Dispersion(og, O.DISP_FIXED, wl, resolution=85_000.0)
    true = np.array([5900.0, -0.3, 3.62, 12.0, 20.0])  # just above the 5750-6250 step
    flux, var = synth.observe(O.produce_model(og, od, true), wl, 100.0, 2100)
    print(f"staircase grid: {len(labels)} models, limits {LIMITS}", flush=True)
    return pasta, synth, spec, wlg, interp, og, wl, od, flux, var


def _stair_labels(og, spec, n, seed):
    """n labels inside the staircase (rejection sampled) plus the steps' corners and edges."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        t, m, g = rng.uniform(5000, 7000), rng.uniform(-1, 0.5), rng.uniform(3.0, 5.0)
        if O._L.or_is_teff_logg_between_bounds(og._h, t, g):
            out.append((t, m, g))
    edges = [(5625.0, 0.1, 3.5), (5750.0, -1.0, 3.5),
             (5760.0, 0.3, 3.51), (6250.0, 0.25, 4.0), (6375.0, -0.6, 4.0),
             (6500.0, 0.0, 4.0), (7000.0, 0.5, 5.0), (6990.0, -0.99, 4.01), (5000.0, -1.0, 3.0),
             (5125.0, 0.45, 4.95), (6100.0, -0.5, 3.55), (6600.0, 0.2, 4.499)]
    return np.array(out + edges)


def test_staircase_weights_and_interpolation(stair):
    pasta, synth, spec, wlg, interp, og, wl, od, flux, var = stair
    labs = _stair_labels(og, spec, 220, 1)
    print(f"staircase interpolate: {len(labs)} labels", flush=True)
    not_one = 0
    for lab in labs:
        idx, w = interp._weights(*lab)
        oidx, ow32, ow64, *_ = og.local_weights(*lab)
        np.testing.assert_array_equal(idx, oidx)
        np.testing.assert_array_equal(w, ow32)
        not_one += abs(float(ow64.sum()) - 1.0) > 1e-9
    assert not_one > 0  # labels where the staircase leaves weights not summing to 1
    out, st = interp.interpolate_bulk(labs)
    assert not st.any()
    for k, lab in enumerate(labs):
        ref = og.interpolate(*lab)
        assert _rel(out[k], ref) < RTOL, (lab.tolist(), _rel(out[k], ref))
    # outside the staircase (below a step), and on the last Teff column before a step at a
    # logg the next column lacks (a grid point, but the logg limits at that Teff take the
    # narrower neighbour's, src/interpolate.rs:378-398): the bounds error on both sides
    bad = np.array([[6600.0, 0.0, 3.2], [5900.0, 0.0, 3.1], [5500.0, 0.0, 3.0],
                    [5500.0, -0.2, 3.2], [6250.0, 0.5, 3.5]])
    out, st = interp.interpolate_bulk(bad)
    assert (st == 1).all(), st
    for lab in bad:
        with pytest.raises(O.OracleError) as e:
            og.interpolate(*lab)
        assert e.value.code == -1


def test_staircase_models_and_chi2(stair):
    pasta, synth, spec, wlg, interp, og, wl, od, flux, var = stair
    rng = np.random.default_rng(2)
    lab3 = _stair_labels(og, spec, 40, 3)
    labs = np.concatenate([lab3, rng.uniform(0, 250, (len(lab3), 1)),
                           rng.uniform(-140, 140, (len(lab3), 1))], axis=1)
    d = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
    models = interp.produce_model_bulk(d, labs)
    for k, lab in enumerate(labs):
        assert _rel(models[k], O.produce_model(og, od, lab)) < RTOL, lab.tolist()
    oc = O.Continuum("chunk", len(wl), wl=wl, n_chunks=5, p_order=4, overlap=0.2)
    star = pasta.Star(interp, d, pasta.ChunkContinuumFitter(wl, 5, 4, 0.2), flux, var)
    chi2, st, _ = star.chi2(labs)
    assert not st.any()
    ref = O.cost_batch(og, od, oc, flux, var, labs, 8)
    rel = np.abs(chi2 - ref) / ref
    print(f"staircase chi2: max rel {rel.max():.2e} over {len(labs)} labels", flush=True)
    assert rel.max() < RTOL


CONSTRAINTS = {
    "none": [],
    # (logg 3.6 with Teff up to 6400 would pin particles to Teff 6250, the Teff limit at
    # logg 3.6, where the logg limits at that Teff exclude 3.6: the reference's cost fails
    # there and aborts the fit, and so do the oracle and the GPU)
    "fix_logg_range_teff": [(2, 1, 4.2, 4.2), (0, 2, 5600.0, 6400.0)],
    "fix_vsini_range_rv": [(3, 1, 12.0, 12.0), (4, 2, 0.0, 60.0)],
}


@pytest.mark.parametrize("name", list(CONSTRAINTS))
def test_staircase_fit_traces_equal_oracle_pso(stair, name):
    """Plain (host PSO) and batched (device PSO) fits on the staircase grid, with and
    without constraints, against the oracle's PSO fed the GPU's chi2: identical traces."""
    pasta, synth, spec, wlg, interp, og, wl, od, flux, var = stair
    cons = CONSTRAINTS[name]
    P, iters, seed = 24, 20, 61
    d = pasta.FixedResolutionDispersion(wl, 85_000.0, wlg)
    c = pasta.ChunkContinuumFitter(wl, 5, 4, 0.2)
    fitter = interp.get_fitter(d, c, pasta.PSOSettings(P, iters), (0.0, 300.0), (-150.0, 150.0))
    pcons = [pasta.FixConstraint(p, a) if k == 1 else pasta.RangeConstraint(p, a, b)
             for (p, k, a, b) in cons]
    print(f"staircase fit [{name}]", flush=True)
    res = fitter.fit(interp, flux, var, seed=seed, constraints=pcons)
    star = pasta.Star(interp, d, c, flux, var)
    b = O.bounds(og, (0.0, 300.0), (-150.0, 150.0), cons)
    best, bc, trace, ne = O.pso_run(b, O.pso_settings(P, iters), seed,
                                    lambda pos: star.chi2(pos)[0])
    assert ne == P * (iters + 1)
    np.testing.assert_array_equal(trace, res._trace)
    assert bc == res.chi2 and np.array_equal(best, res.label.as_list())
    for (p, k, a, bb) in cons:
        assert (k == 1 and trace[-1, 1 + p] == a) or (k == 2 and a <= trace[-1, 1 + p] <= bb)
    bulk = fitter.fit_bulk(interp, [flux], [var], seeds=[seed], trace=True, constraints=pcons)[0]
    np.testing.assert_array_equal(bulk._trace, res._trace)
