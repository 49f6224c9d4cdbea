"""Pins of the oracle FMM (oracle/oracle.c, restating src/eikonal_fmm.cpp:12-71) -- CPU.

Each check is independent of the heap-based FMM code: 1-D strips have the closed form
T = prefix sums of cost; the discrete system's residual must vanish; a Gauss-Seidel
fast-sweeping fixed point (a different algorithm) must agree; obstacles and unreachable
regions stay +INF; symmetry of the grid maps onto the solution.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import fixed_point_gap, ulp32


@pytest.mark.parametrize("fmm", [oracle.fmm_f32, oracle.fmm_f64])
def test_strip_is_prefix_sum(fmm, rng):
    c = rng.uniform(0.1, 5.0, (1, 97)).astype(np.float32)
    t = fmm(c, [(0, 0)])
    expect = np.concatenate([[0.0], np.cumsum(c[0, 1:].astype(np.float64))])
    assert np.allclose(t[0], expect, rtol=1e-5)
    # seed in the middle: sums run both ways
    t = fmm(c, [(40, 0)])
    left = np.concatenate([np.cumsum(c[0, :40][::-1].astype(np.float64))[::-1], [0.0]])
    assert np.allclose(t[0, :41], left, rtol=1e-5)
    # vertical strip behaves the same
    tv = fmm(c.T.copy(), [(0, 40)])
    assert np.allclose(tv[:, 0], t[0], rtol=1e-6)


def test_constant_cost_closed_forms():
    c = np.full((9, 9), 2.0, dtype=np.float64)
    t = oracle.fmm_f64(c, [(0, 0)])
    assert np.allclose(t[0, :], 2.0 * np.arange(9))            # along the axes: k c
    assert np.allclose(t[:, 0], 2.0 * np.arange(9))
    assert t[1, 1] == pytest.approx(2.0 * (1 + math.sqrt(2) / 2))
    # bounds of the first-order upwind scheme: Euclidean <= T <= Manhattan (in units of c)
    yy, xx = np.mgrid[0:9, 0:9]
    assert np.all(t <= 2.0 * (xx + yy) + 1e-9)
    assert np.all(t >= 2.0 * np.sqrt(xx ** 2 + yy ** 2) - 1e-9)
    assert np.allclose(t, t.T)                                  # diagonal symmetry


def test_residual_vanishes(rng):
    c = rng.uniform(0.1, 10.0, (64, 80))
    seeds = [(3, 5), (70, 60)]
    t = oracle.fmm_f64(c, seeds)
    assert np.all(fixed_point_gap(t, c, seeds) <= 1e-11 * np.maximum(t, 1.0))
    t32 = oracle.fmm_f32(c.astype(np.float32), seeds)
    assert np.all(fixed_point_gap(t32, c.astype(np.float32), seeds) <= 4 * ulp32(t32))
    for x, y in seeds:
        assert t[y, x] == 0.0


def test_matches_sweeping_fixed_point(rng):
    for trial in range(4):
        h, w = rng.integers(5, 60, 2)
        c = rng.uniform(0.05, 20.0, (h, w))
        c[rng.random((h, w)) < 0.15] = np.inf                   # obstacles
        seeds = {(int(rng.integers(0, w)), int(rng.integers(0, h))) for _ in range(3)}
        for x, y in seeds:
            c[y, x] = 1.0
        seeds = sorted(seeds)
        a = oracle.fmm_f64(c, seeds)
        b, rounds = oracle.sweep_f64(c, seeds)
        assert rounds < 10000
        fa, fb = np.isfinite(a), np.isfinite(b)
        assert np.array_equal(fa, fb)
        assert np.allclose(a[fa], b[fb], rtol=1e-12, atol=1e-12)


def test_obstacles_and_unreachable():
    c = np.ones((7, 7), dtype=np.float32)
    c[:, 3] = np.inf                                            # full wall splits the grid
    t = oracle.fmm_f32(c, [(0, 0)])
    assert np.all(np.isinf(t[:, 3:]))
    assert np.all(np.isfinite(t[:, :3]))
    t2 = oracle.fmm_f32(c, [(0, 0), (6, 6)])
    assert np.all(np.isinf(t2[:, 3]))
    assert np.isfinite(t2).sum() == 42


def test_f32_close_to_f64(rng):
    c = rng.uniform(0.1, 10.0, (120, 100)).astype(np.float32)
    seeds = [(10, 10), (90, 100)]
    a = oracle.fmm_f32(c, seeds)
    b = oracle.fmm_f64(c.astype(np.float64), seeds)
    m = oracle.error_metrics(a, b)
    assert m["finiteness_mismatch"] == 0
    assert m["linf"] <= 1e-5 * m["max_ref"]


def test_bad_arguments():
    c = np.ones((4, 4), dtype=np.float32)
    with pytest.raises(ValueError):
        oracle.fmm_f32(c, [(4, 0)])


def test_error_metrics_hand_computed():
    got = np.array([0.0, 1.0, 2.5, np.inf, 3.0], dtype=np.float32)
    ref = np.array([0.0, 1.5, 2.0, np.inf, np.inf])
    m = oracle.error_metrics(got, ref)
    assert m["linf"] == 0.5
    assert m["l1"] == pytest.approx(1.0 / 3.0)
    assert m["finiteness_mismatch"] == 1
    assert m["compared"] == 3
    assert m["max_ref"] == 2.0
