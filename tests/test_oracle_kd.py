"""Pins of the K-D oracle (oracle/oracle.c): the generic K-axis update eikonal_solve<K>
(src/eikonal.h:73-118) and the three-dimensional FMM and sweeping solvers -- CPU.

The update is pinned by closed forms (K = 1: a + c; three equal minima: a + c/sqrt3),
by the 2-D forms it must reduce to, by the residual of its defining equation and by
monotonicity.  The 3-D FMM is pinned by 1-D strips (prefix sums), D = 1 slabs (the
pinned 2-D FMM), constant-cost closed forms, the first-order bounds, obstacles and a
Gauss-Seidel fixed point (a different algorithm).
"""
import math

import numpy as np
import pytest

import oracle
from conftest import fixed_point_gap3d, godunov3_f64

INF = math.inf


@pytest.mark.parametrize("gk", [oracle.godunov_k_f32, oracle.godunov_k_f64])
def test_k_update_closed_forms(gk):
    assert gk([2.0], 3.0) == pytest.approx(5.0)
    assert gk([0.0, 0.0], 1.0) == pytest.approx(math.sqrt(2) / 2, rel=1e-6)
    assert gk([0.0, 0.0, 0.0], 1.0) == pytest.approx(1 / math.sqrt(3), rel=1e-6)
    assert gk([4.0, 4.0, 4.0], 3.0) == pytest.approx(4.0 + 3.0 / math.sqrt(3), rel=1e-6)
    # far axes are ignored: (0, 5, 9) with cost 1 -> 1
    assert gk([9.0, 0.0, 5.0], 1.0) == pytest.approx(1.0)
    # boundary between one and two axes: b = a + c -> a + c
    assert gk([0.0, 1.0], 1.0) == pytest.approx(1.0)
    # INF axes
    assert gk([INF, INF, INF], 1.0) == INF
    assert gk([1.0, INF, INF], 2.0) == pytest.approx(3.0)
    # min against the previous best
    assert gk([0.0, 0.0, 0.0], 1.0, 0.25) == pytest.approx(0.25)


def test_k_update_reduces_to_2d(rng):
    for _ in range(2000):
        a, b = rng.uniform(0, 50, 2)
        c = rng.uniform(0.01, 20)
        t2 = oracle.godunov_f64(a, b, c)
        assert oracle.godunov_k_f64([a, b], c) == pytest.approx(t2, rel=1e-12, abs=1e-12)
        assert oracle.godunov_k_f64([a, INF, b], c) == pytest.approx(t2, rel=1e-12, abs=1e-12)


def test_k_update_residual_and_monotone(rng):
    for _ in range(3000):
        m = rng.uniform(0, 30, 3)
        c = rng.uniform(0.01, 20)
        t = oracle.godunov_k_f64(m, c)
        res = sum(max(t - x, 0.0) ** 2 for x in m)
        assert res == pytest.approx(c * c, rel=1e-9)
        assert t >= m.min()
        for k in range(3):
            up = m.copy()
            up[k] += 0.25
            assert oracle.godunov_k_f64(up, c) >= t - 1e-12
        assert oracle.godunov_k_f64(m, c * 1.1) >= t
        # the FP32 form stays within a few ulp-scale errors of FP64 at these magnitudes
        t32 = oracle.godunov_k_f32(m.astype(np.float32), np.float32(c))
        assert abs(t32 - t) <= 1e-4 * max(t, 1.0)


def test_k_update_matches_definition_helper(rng):
    """The test helper godunov3_f64 (used for 3-D residuals) agrees with the oracle."""
    m = rng.uniform(0, 30, (3, 500))
    c = rng.uniform(0.01, 20, 500)
    g = godunov3_f64(m[0], m[1], m[2], c)
    o = np.array([oracle.godunov_k_f64(m[:, i], c[i]) for i in range(500)])
    assert np.allclose(g, o, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("fmm", [oracle.fmm3d_f32, oracle.fmm3d_f64])
def test_fmm3d_strips_are_prefix_sums(fmm, rng):
    c = rng.uniform(0.1, 5.0, 61)
    for shape, seed in (((1, 1, 61), (0, 0, 0)), ((1, 61, 1), (0, 0, 0)), ((61, 1, 1), (0, 0, 0))):
        t = fmm(c.reshape(shape), [seed]).ravel()
        expect = np.concatenate([[0.0], np.cumsum(c[1:])])
        assert np.allclose(t, expect, rtol=1e-5)


def test_fmm3d_slab_equals_2d(rng):
    c = rng.uniform(0.1, 10.0, (1, 37, 45))
    seeds2 = [(3, 4), (40, 30)]
    t3 = oracle.fmm3d_f64(c, [(x, y, 0) for x, y in seeds2])
    t2 = oracle.fmm_f64(c[0], seeds2)
    assert np.allclose(t3[0], t2, rtol=1e-12, atol=1e-12)


def test_fmm3d_constant_cost_closed_forms():
    c = np.full((7, 7, 7), 2.0)
    t = oracle.fmm3d_f64(c, [(0, 0, 0)])
    v = 2.0 * (1 + math.sqrt(2) / 2)
    assert np.allclose(t[0, 0, :], 2.0 * np.arange(7))
    assert np.allclose(t[:, 0, 0], 2.0 * np.arange(7))
    assert t[0, 1, 1] == pytest.approx(v)
    assert t[1, 1, 1] == pytest.approx(v + 2.0 / math.sqrt(3))
    zz, yy, xx = np.mgrid[0:7, 0:7, 0:7]
    assert np.all(t <= 2.0 * (xx + yy + zz) + 1e-9)
    assert np.all(t >= 2.0 * np.sqrt(xx ** 2 + yy ** 2 + zz ** 2) - 1e-9)
    # the three axes are interchangeable
    assert np.allclose(t, np.transpose(t, (2, 1, 0)))
    assert np.allclose(t, np.transpose(t, (1, 0, 2)))


def test_fmm3d_equals_sweeping_fixed_point(rng):
    c = rng.uniform(0.1, 10.0, (9, 11, 13))
    c[4, 2:9, 3:10] = np.inf                    # a wall with gaps around it
    seeds = [(1, 1, 1), (12, 9, 7)]
    t = oracle.fmm3d_f64(c, seeds)
    s, rounds = oracle.sweep3d_f64(c, seeds)
    assert rounds < 10000
    fin = np.isfinite(t)
    assert np.array_equal(fin, np.isfinite(s))
    assert np.allclose(t[fin], s[fin], rtol=1e-12, atol=1e-12)
    assert np.all(fixed_point_gap3d(t, c, seeds) <= 1e-10 * np.maximum(t, 1.0))


def test_fmm3d_f32_close_to_f64(rng):
    c = rng.uniform(0.5, 2.0, (12, 14, 16))
    seeds = [(0, 0, 0), (15, 13, 11)]
    t64 = oracle.fmm3d_f64(c, seeds)
    t32 = oracle.fmm3d_f32(c.astype(np.float32), seeds)
    assert np.max(np.abs(t32 - t64)) <= 1e-4 * t64.max()


def test_fmm3d_obstacles_and_unreachable():
    c = np.ones((6, 6, 6))
    c[3, :, :] = np.inf                          # a full plane cuts the box in two
    t = oracle.fmm3d_f64(c, [(0, 0, 0)])
    assert np.all(np.isinf(t[3:]))
    assert np.all(np.isfinite(t[:3]))
    with pytest.raises(ValueError):
        oracle.fmm3d_f64(c, [(6, 0, 0)])
