"""compute_uncertainty in the oracle (src/fitting.rs:333-372,482-568).  Its root finder
restates argmin 0.10.0's BrentRoot (absent from the reference tree); both are pinned by
closed forms: known roots, the bracket-sign error, and - for the uncertainty itself -
that chi2 - target changes sign across every reported crossing."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("fn,lo,hi,root", [
    (lambda x: x * x - 0.25, 0.0, 1.0, 0.5),
    (lambda x: np.cos(x) - x, 0.0, 1.0, 0.7390851332151607),
    (lambda x: x ** 3 - 2 * x - 5, 2.0, 3.0, 2.0945514815423265),
    (lambda x: -(x + 0.3) * (x + 3.0), -1.0, 0.0, -0.3),      # left bracket, decreasing
    (lambda x: np.tanh(40.0 * (x - 0.123)), 0.0, 1.0, 0.123),  # steep
])
def test_brent_roots(fn, lo, hi, root):
    for tol in (1e-3, 1e-8):
        r, n = O.brent_root(fn, lo, hi, tol)
        assert abs(r - root) <= tol, (r, root, tol)
        assert n < 60


def test_brent_same_sign_is_an_error():
    with pytest.raises(O.OracleError):
        O.brent_root(lambda x: x * x + 1.0, 0.0, 1.0, 1e-3)


def test_brent_root_at_bracket_end():
    r, _ = O.brent_root(lambda x: x - 1.0, 0.0, 1.0, 1e-3)
    assert r == 1.0


@pytest.mark.parametrize("radius", [None, (0.01, 0.1, 0.1, 1.0, 1.0)])
def test_uncertainty_crossings(small, radius):
    spec_res = 85_000.0
    label = small.true
    out, st = O.compute_uncertainty(small.og, small.odisp, small.ocont, small.flux, small.var,
                                    spec_res, label, radius)
    best, _ = O.cost(small.og, small.odisp, small.ocont, small.flux, small.var, label)
    wl = small.wl
    n = 4.0 * spec_res * (wl[-1] - wl[0]) / (wl[0] + wl[-1])
    target = best * (1.0 + np.sqrt(2.0 / n))
    sr = np.array([2000.0, 0.3, 0.3, 40.0, 40.0] if radius is None else radius)
    sr[0] *= label[0]
    sr[3] = max(label[3] * sr[3], 100.0)
    found = checked = 0
    for i in range(5):
        for k, sgn in ((0, -1.0), (1, 1.0)):
            if st[i, k]:
                continue
            found += 1
            d = out[i, k]
            assert d >= 0.0
            x = sgn * d / sr[i]           # rescaled crossing
            lab = label.copy()

            def f(xx):
                lab[i] = label[i] + xx * sr[i]
                return O.cost(small.og, small.odisp, small.ocont, small.flux, small.var, lab)[0] - target
            # Brent's tolerance is 1e-3 of the radius: chi2 - target changes sign within it
            try:
                vals = [f(x - 1e-3), f(x), f(x + 1e-3)]
            except O.OracleError:   # the check point left the grid (teff's radius is huge)
                continue
            checked += 1
            assert min(vals) <= 0.0 <= max(vals), (i, k, d, vals)
    assert found >= 8 and checked >= 6, (st, checked)
