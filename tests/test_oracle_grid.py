"""Pins the oracle's grid geometry, neighbour choice and Lagrange weights
(reference src/interpolate.rs:58-533, src/cubic.rs) against what the mathematics
fixes: polynomial reproduction, partition of unity, one-hot weights at nodes,
brute-force neighbour invariants on staircase grids, and the storage index."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def staircase_labels(teff, m, logg, limits):
    out = []
    for i, t in enumerate(teff):
        lo, hi = limits[i]
        for k in range(lo, hi + 1):
            for mm in m:
                out.append((t, mm, logg[k]))
    return np.array(out)


TEFF = np.arange(5000.0, 7001.0, 250.0)
MH = np.array([-1.0, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5])
LOGG = np.arange(3.0, 5.01, 0.5)
STAIR = [(0, 4), (0, 4), (0, 4), (1, 4), (1, 4), (1, 3), (2, 3), (2, 3), (2, 3)]


@pytest.fixture(scope="module")
def rect():
    return O.Grid(staircase_labels(TEFF, MH, LOGG, [(0, 4)] * len(TEFF)))


@pytest.fixture(scope="module")
def stair():
    return O.Grid(staircase_labels(TEFF, MH, LOGG, STAIR))


def test_axes_and_storage_order(stair):
    t, m, g, lo, hi, cum = stair.axes()
    assert np.array_equal(t, TEFF) and np.array_equal(m, MH) and np.array_equal(g, LOGG)
    assert [(a, b) for a, b in zip(lo, hi)] == STAIR
    pts = stair.list_gridpoints()
    # list_gridpoints order is the order the index formula addresses
    for q, (tt, mm, gg) in enumerate(pts):
        i = int(np.where(TEFF == tt)[0][0]); j = int(np.where(MH == mm)[0][0])
        k = int(np.where(LOGG == gg)[0][0])
        assert O._L.or_grid_model_index(stair._h, i, j, k) == q


def _poly(t, m, g, deg_t=3, deg_m=3, deg_g=3):
    # product of one polynomial per axis, degree <= the given ones
    x = (t - 6000.0) / 1000.0
    y = m + 0.25
    z = g - 4.0
    return ((1 + 0.3 * x - 0.2 * x ** 2 + 0.1 * x ** 3 * (deg_t >= 3))
            * (2 - y + 0.5 * y ** 2 + 0.25 * y ** 3 * (deg_m >= 3))
            * (1 + z + 0.3 * z ** 2 - 0.2 * z ** 3 * (deg_g >= 3)))


def _node_values(grid, idx):
    pts = grid.list_gridpoints()
    return np.array([_poly(*pts[r]) if r >= 0 else 0.0 for r in idx])


@pytest.mark.parametrize("seed", range(5))
def test_cubic_reproduction_interior(rect, seed):
    rng = np.random.default_rng(seed)
    for _ in range(40):
        t = rng.uniform(5600, 6400); m = rng.uniform(-0.6, 0.1); g = rng.uniform(3.6, 4.4)
        idx, w32, w64, *_ = rect.local_weights(t, m, g)
        assert (idx >= 0).all()
        assert abs(w64.sum() - 1.0) < 1e-12
        val = np.dot(w64, _node_values(rect, idx))
        assert abs(val - _poly(t, m, g)) < 1e-9 * abs(_poly(t, m, g))
        assert np.array_equal(w32, w64.astype(np.float32))


def test_quadratic_reproduction_at_edges(rect):
    rng = np.random.default_rng(1)
    for _ in range(60):
        t = rng.choice([rng.uniform(5000, 5250), rng.uniform(6750, 7000)])
        m = rng.choice([rng.uniform(-1.0, -0.75), rng.uniform(0.25, 0.5)])
        g = rng.choice([rng.uniform(3.0, 3.5), rng.uniform(4.5, 5.0)])
        idx, _, w64, *_ = rect.local_weights(t, m, g)
        assert abs(w64.sum() - 1.0) < 1e-12
        pts = rect.list_gridpoints()
        vals = np.array([_poly(*pts[r], 2, 2, 2) if r >= 0 else 0.0 for r in idx])
        exact = _poly(t, m, g, 2, 2, 2)
        assert abs(np.dot(w64, vals) - exact) < 1e-9 * abs(exact)


def test_one_hot_at_interior_nodes(rect):
    pts = rect.list_gridpoints()
    for (t, m, g) in pts[::7]:
        idx, _, w64, *_ = rect.local_weights(t, m, g)
        nz = np.abs(w64) > 1e-12
        assert nz.sum() == 1
        assert abs(w64[nz][0] - 1.0) < 1e-12
        assert tuple(pts[idx[nz][0]]) == (t, m, g)


def _brute_neighbors(values, x, left, right):
    """Independent statement of the neighbour rule: the 4 consecutive indices
    around x, dropping the one that falls outside [left, right]."""
    n = len(values)
    if x < values[left] or x > values[right]:
        return None
    # index of the grid interval [a, a+1] containing x (exact nodes: the right interval,
    # except at the right limit)
    a = max(left, min(right - 1, int(np.searchsorted(values, x, side="right")) - 1))
    cand = [a - 1, a, a + 1, a + 2]
    return [c if left <= c <= right else -1 for c in cand]


def test_find_neighbors_brute_force():
    vals = np.arange(10.0)
    out = np.zeros(4, np.int32)
    vp = vals.ctypes.data_as(O.C.POINTER(O.C.c_double))
    rng = np.random.default_rng(3)
    for left, right in [(0, 9), (2, 7), (1, 4)]:
        xs = np.concatenate([rng.uniform(left, right, 50), np.arange(left, right + 1.0)])
        for x in xs:
            rc = O._L.or_find_neighbors(vp, 10, float(x), left, right,
                                        out.ctypes.data_as(O.C.POINTER(O.C.c_int)))
            exp = _brute_neighbors(vals, x, left, right)
            assert rc == 0
            got = out.tolist()
            present = [c for c in got if c >= 0]
            # consecutive, inside the limits, bracketing x
            assert present == list(range(present[0], present[0] + len(present)))
            assert all(left <= c <= right for c in present)
            assert vals[got[1]] <= x <= vals[got[2]] or (x == vals[got[1]])
            # only dropped when the 4-window would leave the limits
            assert len(present) >= 3
            if len(present) == 3:
                assert got[0] == -1 and got[1] == left or got[3] == -1 and got[2] == right
            assert set(present) <= set(c for c in exp if c >= 0) | set(present)
        for x in (left - 0.5, right + 0.5):
            rc = O._L.or_find_neighbors(vp, 10, float(x), left, right,
                                        out.ctypes.data_as(O.C.POINTER(O.C.c_int)))
            assert rc != 0


def test_staircase_bounds_and_weights(stair):
    rng = np.random.default_rng(5)
    n_ok = 0
    for _ in range(400):
        t = rng.uniform(5000, 7000); m = rng.uniform(-1, 0.5); g = rng.uniform(3.0, 5.0)
        inside = O._L.or_is_teff_logg_between_bounds(stair._h, t, g)
        try:
            idx, _, w64, *_ = stair.local_weights(t, m, g)
        except O.OracleError:
            continue
        assert inside
        n_ok += 1
        # The weight of neighbour (r, c, mi) is FT[r,c] * FL[r,c] * FM[mi] (src/cubic.rs:237-241),
        # FT: teff Lagrange factors per logg row over the teff nodes present in that row,
        # FL: logg factors per teff column.  Each present row / column is a Lagrange basis
        # (sums to 1); in staircase regions the rows differ, so the 64 weights need not sum to 1.
        _, _, _, ft, fl, fm = stair.local_weights(t, m, g)
        assert abs(fm.sum() - 1.0) < 1e-12
        for r in range(4):
            if np.any(ft[r] != 0):
                assert abs(ft[r].sum() - 1.0) < 1e-9
        for c in range(4):
            if np.any(fl[:, c] != 0):
                assert abs(fl[:, c].sum() - 1.0) < 1e-9
        prod = np.array([ft[r, c] * fl[r, c] * fm[mi] for c in range(4) for r in range(4)
                         for mi in range(4)])
        np.testing.assert_array_equal(prod, w64)
        pts = stair.list_gridpoints()
        # every used neighbour is a real grid point of the staircase
        for r, w in zip(idx, w64):
            if abs(w) > 0:
                assert r >= 0
                tt, mm, gg = pts[r]
                i = int(np.where(TEFF == tt)[0][0])
                k = int(np.where(LOGG == gg)[0][0])
                assert STAIR[i][0] <= k <= STAIR[i][1]
    assert n_ok > 100


def test_out_of_bounds_errors(rect):
    for lab in [(4999.0, 0.0, 4.0), (6000.0, 0.6, 4.0), (6000.0, 0.0, 2.9), (7001.0, 0.0, 4.0)]:
        with pytest.raises(O.OracleError):
            rect.local_weights(*lab)


def test_interpolation_reproduces_linear_grid_with_tail_quirk():
    """A grid whose integer flux is linear in the labels is reproduced exactly by
    the cubic weights; the last W % 128 pixels carry the reference's extra 1/65535
    (src/interpolate.rs:962 vs :978)."""
    labels = staircase_labels(TEFF, MH, LOGG, [(0, 4)] * len(TEFF))
    W = 300
    g = O.Grid(labels)
    pts = g.list_gridpoints()
    ti = (pts[:, 0] - 5000) / 250; mi = (pts[:, 1] + 1) / 0.25; gi = (pts[:, 2] - 3) / 0.5
    px = np.arange(W)
    spectra = (1000 + 50 * ti[:, None] + 30 * mi[:, None] + 20 * gi[:, None] + px[None, :]).astype(np.uint16)
    g = O.Grid(labels, spectra, None, 3.7, 1e-5)
    for (t, m, lg) in [(5600.0, -0.3, 3.8), (6100.5, 0.1, 4.6), (5000.0, -1.0, 3.0)]:
        exp = 1000 + 50 * (t - 5000) / 250 + 30 * (m + 1) / 0.25 + 20 * (lg - 3) / 0.5 + px
        out64 = g.interpolate(t, m, lg, f64=True)
        out32 = g.interpolate(t, m, lg)
        nb = (W // 128) * 128
        np.testing.assert_allclose(out64[:nb] * 65535, exp[:nb], rtol=1e-12)
        np.testing.assert_allclose(out64[nb:] * 65535 * 65535, exp[nb:], rtol=1e-12)
        np.testing.assert_allclose(out32[:nb] * 65535, exp[:nb], rtol=2e-6)
        np.testing.assert_allclose(out32[nb:] * 65535.0 * 65535.0, exp[nb:], rtol=2e-6)


def test_factors_scale_each_model():
    labels = staircase_labels(TEFF[:4], MH[:4], LOGG[:4], [(0, 3)] * 4)
    g0 = O.Grid(labels)
    n = g0.n_models
    rng = np.random.default_rng(2)
    spectra = rng.integers(1000, 60000, (n, 256)).astype(np.uint16)
    fac = (rng.uniform(1e3, 1e7, n).astype(np.float32) / np.float32(65535))
    g = O.Grid(labels, spectra, fac, 3.7, 1e-5)
    idx, w32, w64, *_ = g.local_weights(5400.0, -0.6, 3.9)
    out = g.interpolate(5400.0, -0.6, 3.9, f64=True)
    exp = sum(w64[q] * fac[idx[q]] * spectra[idx[q]].astype(np.float64) for q in range(64) if idx[q] >= 0)
    np.testing.assert_allclose(out[:128], exp[:128], rtol=1e-12)


def test_neighbour_index_outside_the_axis_is_an_error():
    """Where an axis' limits span two points, the neighbour rules name an index outside the
    axis (i + 2 past the end at the left limit, i - 2 below zero at the right, or at the
    middle node of three, src/interpolate.rs:124-132): the reference's value lookup then
    panics (:483-530; i - 2 wraps first), the oracle returns "not enough neighbours".
    Four points are enough everywhere."""
    def grid(ms):
        return O.Grid(staircase_labels(TEFF[:4], np.array(ms), LOGG, [(0, 4)] * 4))

    def ok(g, m):
        idx, w32, w64, *_ = g.local_weights(5300.0, m, 4.1)
        assert abs(w64.sum() - 1.0) < 1e-12 and (idx < g.n_models).all()

    def err(g, m):
        with pytest.raises(O.OracleError) as e:
            g.local_weights(5300.0, m, 4.1)
        assert e.value.code == -2, m

    two = grid([-0.5, 0.0])
    for m in (-0.5, -0.2, 0.0):
        err(two, m)
    three = grid([-0.5, 0.0, 0.5])
    for m in (-0.5, -0.2, 0.3, 0.5):
        ok(three, m)
    err(three, 0.0)  # exact middle node: i == right - 1 = 1, neighbour i - 2 = -1
    four = grid([-0.5, 0.0, 0.5, 1.0])
    for m in (-0.5, -0.2, 0.0, 0.3, 0.5, 0.7, 1.0):
        ok(four, m)
