"""Pins of the oracle FIM (restating src/eikonal_fim.cpp:13-107) -- CPU."""
import numpy as np
import pytest

import oracle


def test_fim_matches_fmm(rng):
    for eps in (1e-5, 1e-3, 0.0):
        c = rng.uniform(0.1, 10.0, (70, 90)).astype(np.float32)
        c[20:50, 40] = np.inf
        seeds = [(5, 5), (80, 60)]
        a = oracle.fmm_f32(c, seeds)
        b, solves = oracle.fim_f32(c, seeds, eps)
        fa = np.isfinite(a)
        assert np.array_equal(fa, np.isfinite(b))
        assert np.max(np.abs(a[fa] - b[fa])) <= max(1e-4 * a[fa].max(), 50 * eps)
        assert solves >= fa.sum() - len(seeds)


def test_fim_seed_adjacency_quirk():
    # Two 4-adjacent seeds: each is queued as the other's neighbour and solved with the
    # CPU form, whose INF branch skips the min (src/eikonal.h:127-128).  On this 3x4 grid
    # the second seed processed is raised from 0 to cost + 0 = 3 and the cells to its
    # right follow from that value; FMM keeps both seeds at 0.  (On a 1-cell-high strip
    # the same quirk makes the reference FIM climb until FP32 stops changing.)
    c = np.full((3, 4), 3.0, dtype=np.float32)
    seeds = [(1, 1), (2, 1)]
    t, _ = oracle.fim_f32(c, seeds, 1e-5)
    f = oracle.fmm_f32(c, seeds)
    assert f[1, 1] == 0.0 and f[1, 2] == 0.0
    assert t[1, 1] == 0.0 and t[1, 2] == 3.0
    assert t[1, 3] == 6.0 and f[1, 3] == 3.0
    assert np.array_equal(t[:, :2], f[:, :2])


def test_fim_single_cell_and_strip():
    c = np.full((1, 1), 1.0, dtype=np.float32)
    t, n = oracle.fim_f32(c, [(0, 0)])
    assert t[0, 0] == 0.0 and n == 0
    c = np.full((1, 6), 0.5, dtype=np.float32)
    t, _ = oracle.fim_f32(c, [(0, 0)])
    assert np.allclose(t[0], 0.5 * np.arange(6))
