"""Pins the oracle's PSO (reference src/particleswarm.rs, src/bounds.rs) by its
invariants: every particle stays inside the (staircase, constrained) bounds, the
global best never increases, P*(iters+1) evaluations, injected populations are
used, a quadratic bowl is solved, fixed constraints hold exactly, and the
counter-based stream is reproducible."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_grid import LOGG, MH, STAIR, TEFF, staircase_labels


@pytest.fixture(scope="module")
def grid():
    return O.Grid(staircase_labels(TEFF, MH, LOGG, STAIR))


def _bowl(center, scale):
    def f(pos):
        return (((pos - center) / scale) ** 2).sum(axis=1)
    return f


def test_bowl_converges_and_invariants(grid):
    b = O.bounds(grid, (0.0, 300.0), (-150.0, 150.0))
    s = O.pso_settings(40, 100)
    center = np.array([5800.0, -0.2, 4.1, 50.0, 12.0])
    scale = np.array([2000.0, 1.5, 2.0, 300.0, 300.0])
    seen = []

    def f(pos):
        for p in pos:
            assert O._L.or_bounds_within(O.C.byref(b), p.ctypes.data_as(O.C.POINTER(O.C.c_double)))
        seen.append(len(pos))
        return _bowl(center, scale)(pos)

    best, bc, trace, ne = O.pso_run(b, s, 1234, f)
    assert ne == 40 * 101 and sum(seen) == ne
    assert np.all(np.diff(trace[:, 0]) <= 0)
    assert bc == trace[-1, 0]
    assert np.all(np.abs(best - center) <= 1e-3 * scale), (best, center)


def test_reproducible_and_seed_dependent(grid):
    b = O.bounds(grid, (0.0, 300.0), (-150.0, 150.0))
    s = O.pso_settings(16, 20)
    f = _bowl(np.array([6000.0, 0.0, 4.0, 10.0, 0.0]), np.array([1000.0, 1.0, 1.0, 100.0, 100.0]))
    a = O.pso_run(b, s, 7, f)
    bb = O.pso_run(b, s, 7, f)
    c = O.pso_run(b, s, 8, f)
    assert np.array_equal(a[2], bb[2])
    assert not np.array_equal(a[2], c[2])


def test_init_positions_inside_staircase(grid):
    b = O.bounds(grid, (0.0, 300.0), (-150.0, 150.0))
    pos, vel = O.pso_init_positions(b, 500, 99)
    lo = np.array([TEFF[0], MH[0], LOGG[0], 0.0, -150.0])
    hi = np.array([TEFF[-1], MH[-1], LOGG[-1], 300.0, 150.0])
    assert np.all(pos >= lo) and np.all(pos <= hi)
    for p in pos:
        assert O._L.or_is_teff_logg_between_bounds(grid._h, p[0], p[2])
    assert np.all(np.abs(vel) <= (hi - lo))
    # uniform draws: the mean of each free dimension is near the centre of its range
    assert abs(np.mean(pos[:, 3]) - 150.0) < 25.0


def test_population_injection(grid):
    b = O.bounds(grid, (0.0, 300.0), (-150.0, 150.0))
    s = O.pso_settings(8, 0)
    init = np.tile([6000.0, 0.0, 4.0, 20.0, 5.0], (8, 1))
    init[:, 0] += np.arange(8) * 10
    f = _bowl(np.array([6030.0, 0.0, 4.0, 20.0, 5.0]), np.ones(5))
    best, bc, trace, ne = O.pso_run(b, s, 1, f, init_pos=init)
    assert ne == 8
    np.testing.assert_array_equal(best, init[3])


def test_fixed_and_range_constraints(grid):
    b = O.bounds(grid, (0.0, 300.0), (-150.0, 150.0),
                 constraints=[(3, 1, 42.0, 42.0), (4, 2, -10.0, 10.0)])
    s = O.pso_settings(12, 15)
    center = np.array([5800.0, -0.2, 4.1, 100.0, 50.0])
    traj = []

    def f(pos):
        traj.append(pos.copy())
        return _bowl(center, np.array([1000.0, 1.0, 1.0, 100.0, 100.0]))(pos)

    best, *_ = O.pso_run(b, s, 3, f)
    allpos = np.concatenate(traj)
    assert np.all(allpos[:, 3] == 42.0)
    assert np.all((allpos[:, 4] >= -10.0) & (allpos[:, 4] <= 10.0))
    assert best[4] == pytest.approx(10.0)


def test_rng_stream():
    # counter-based SplitMix64: output n of the SplitMix64 sequence seeded with the key
    def splitmix(seed, n):
        z = (seed + (n + 1) * 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return z ^ (z >> 31)
    # SplitMix64 reference output for seed 1234567 (Steele, Lea, Flood 2014; Vigna's splitmix64.c)
    x = 1234567
    seq = []
    for _ in range(3):
        x = (x + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        seq.append(z ^ (z >> 31))
    assert seq[0] == 6457827717110365317
    for n in range(3):
        assert O.rng_u64(1234567, n) == seq[n] == splitmix(1234567, n)
    u = np.array([O.rng_uniform(5, n) for n in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
