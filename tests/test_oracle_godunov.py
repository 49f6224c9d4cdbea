"""Pins of the oracle's Godunov updates (oracle/oracle.c) -- CPU only.

Closed-form known answers, the defining residual, monotonicity, the reference CPU form's
INF quirk (src/eikonal.h:127-128), and agreement between the naive and the
cancellation-safe forms.
"""
import math

import numpy as np
import pytest

import oracle

INF = math.inf
FORMS = [oracle.godunov_cpu, oracle.godunov_gpu, oracle.godunov_safe, oracle.godunov_f64]


@pytest.mark.parametrize("f", FORMS)
def test_closed_forms(f):
    # one finite axis: 1-D update a + c
    assert f(0.0, INF, 1.0) == 1.0
    assert f(INF, 0.0, 1.0) == 1.0
    assert f(2.5, INF, 0.5) == 3.0
    # equal axes: (a + b + sqrt(2) c)/2 = a + c/sqrt(2)
    assert f(0.0, 0.0, 1.0) == pytest.approx(1.0 / math.sqrt(2.0), rel=1e-7)
    assert f(3.0, 3.0, 2.0) == pytest.approx(3.0 + 2.0 / math.sqrt(2.0), rel=1e-7)
    # |a - b| = c is the boundary of both branches: a + c
    assert f(0.0, 1.0, 1.0) == pytest.approx(1.0, rel=1e-7)
    # |a - b| > c: 1-D update from the smaller axis
    assert f(0.0, 2.0, 1.0) == 1.0
    assert f(5.0, 1.0, 0.25) == 1.25
    # nothing known
    assert f(INF, INF, 1.0) == INF
    # obstacle cost stays at the current value
    assert f(1.0, 2.0, INF, 7.0) == 7.0


@pytest.mark.parametrize("f", [oracle.godunov_gpu, oracle.godunov_safe, oracle.godunov_f64])
def test_min_with_current(f):
    assert f(0.0, 0.0, 1.0, 0.5) == 0.5
    assert f(0.0, INF, 1.0, 0.25) == 0.25          # always min'ed


def test_cpu_form_inf_quirk():
    # Reference CPU form returns cost + a without min when the other axis is INF.
    assert oracle.godunov_cpu(0.0, INF, 1.0, 0.25) == 1.0
    assert oracle.godunov_gpu(0.0, INF, 1.0, 0.25) == 0.25


def test_residual_and_monotonicity(rng):
    for _ in range(2000):
        a, b = rng.uniform(0, 100, 2)
        c = float(rng.uniform(1e-3, 20))
        t = oracle.godunov_f64(a, b, c)
        # root of the upwind quadratic: sum of squared positive parts equals c^2
        lhs = max(t - a, 0) ** 2 + max(t - b, 0) ** 2
        assert lhs == pytest.approx(c * c, rel=1e-9)
        assert t > min(a, b)
        # monotone non-decreasing in a, b and c
        assert oracle.godunov_f64(a + 0.5, b, c) >= t
        assert oracle.godunov_f64(a, b + 0.5, c) >= t
        assert oracle.godunov_f64(a, b, c * 1.1) >= t


def _ulp_diff(x: float, y: float) -> int:
    xi = np.array([x], dtype=np.float32).view(np.int32)[0]
    yi = np.array([y], dtype=np.float32).view(np.int32)[0]
    return abs(int(xi) - int(yi))


def test_safe_form_agrees_with_naive(rng):
    worst = 0
    for _ in range(5000):
        a, b = (float(v) for v in rng.uniform(0, 1000, 2).astype(np.float32))
        c = float(np.float32(rng.uniform(0.05, 20)))
        n = oracle.godunov_gpu(a, b, c)
        s = oracle.godunov_safe(a, b, c)
        worst = max(worst, _ulp_diff(n, s))
    assert worst <= 2


def test_f32_forms_close_to_f64(rng):
    for _ in range(2000):
        a, b = (float(v) for v in rng.uniform(0, 100, 2).astype(np.float32))
        c = float(np.float32(rng.uniform(0.01, 10)))
        ref = oracle.godunov_f64(a, b, c)
        for f in (oracle.godunov_cpu, oracle.godunov_gpu, oracle.godunov_safe):
            assert f(a, b, c) == pytest.approx(ref, rel=3e-6, abs=1e-6)
