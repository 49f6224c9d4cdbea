"""Shared pytest setup: markers, repo path, and helpers used by several test files."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def fixed_point_gap(t: np.ndarray, cost: np.ndarray, seeds) -> np.ndarray:
    """|T - G(a, b, cost)| at every non-seed finite cell, G evaluated in FP64.

    The discrete eikonal solution is the fixed point T = G(a, b, c) of the upwind
    Godunov update, with a = min(T[x-1], T[x+1]), b = min(T[y-1], T[y+1]) (out of grid
    = +INF) and G the larger root of max(T-a,0)^2 + max(T-b,0)^2 = c^2.  This is the
    definition of the discrete system, independent of any solver.  Returns 0 at seeds,
    obstacles and unreached cells.
    """
    T = np.asarray(t, dtype=np.float64)
    C = np.asarray(cost, dtype=np.float64)
    P = np.pad(T, 1, constant_values=np.inf)
    a = np.minimum(P[1:-1, :-2], P[1:-1, 2:])
    b = np.minimum(P[:-2, 1:-1], P[2:, 1:-1])
    m = np.minimum(a, b)
    M = np.maximum(a, b)
    with np.errstate(invalid="ignore", over="ignore"):
        d = M - m
        two_axis = (m + M + np.sqrt(np.maximum(2.0 * C * C - d * d, 0.0))) / 2.0
        G = np.where(np.isfinite(M) & (C >= d), two_axis, m + C)
        gap = np.abs(T - G)
    gap[~np.isfinite(T) | ~np.isfinite(C)] = 0.0
    for x, y in np.asarray(seeds).reshape(-1, 2):
        gap[y, x] = 0.0
    return gap


def godunov3_f64(a, b, c, cost):
    """The K = 3 discrete update in FP64 on arrays: the larger root of
    sum_i max(T - m_i, 0)^2 = cost^2 over the sorted axis minima m_0 <= m_1 <= m_2, i.e. the
    first of m_0 + cost (if <= m_1), the two-axis root (if <= m_2), the three-axis root."""
    m = np.sort(np.stack([np.asarray(a, np.float64), np.asarray(b, np.float64),
                          np.asarray(c, np.float64)]), axis=0)
    C = np.asarray(cost, dtype=np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        t1 = m[0] + C
        d = m[1] - m[0]
        two = (m[0] + m[1] + np.sqrt(np.maximum(2.0 * C * C - d * d, 0.0))) / 2.0
        s = m[0] + m[1] + m[2]
        q = s * s - 3.0 * (m[0] ** 2 + m[1] ** 2 + m[2] ** 2 - C * C)
        three = (s + np.sqrt(np.maximum(q, 0.0))) / 3.0
        return np.where(t1 <= m[1], t1, np.where(two <= m[2], two, three))


def fixed_point_gap3d(t: np.ndarray, cost: np.ndarray, seeds) -> np.ndarray:
    """|T - G3(a, b, c, cost)| at every non-seed finite cell of a (D, H, W) field, with the
    axis minima of the six neighbours (out of grid = +INF) and G3 as in godunov3_f64."""
    T = np.asarray(t, dtype=np.float64)
    P = np.pad(T, 1, constant_values=np.inf)
    a = np.minimum(P[1:-1, 1:-1, :-2], P[1:-1, 1:-1, 2:])
    b = np.minimum(P[1:-1, :-2, 1:-1], P[1:-1, 2:, 1:-1])
    c = np.minimum(P[:-2, 1:-1, 1:-1], P[2:, 1:-1, 1:-1])
    G = godunov3_f64(a, b, c, cost)
    with np.errstate(invalid="ignore"):
        gap = np.abs(T - G)
    gap[~np.isfinite(T) | ~np.isfinite(np.asarray(cost, dtype=np.float64))] = 0.0
    for x, y, z in np.asarray(seeds).reshape(-1, 3):
        gap[z, y, x] = 0.0
    return gap


def ulp32(t: np.ndarray) -> np.ndarray:
    """Spacing of FP32 numbers at |t| (finite entries)."""
    a = np.abs(np.asarray(t, dtype=np.float32))
    a = np.where(np.isfinite(a), a, np.float32(1.0))
    return np.spacing(a).astype(np.float64)


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(20261202)
