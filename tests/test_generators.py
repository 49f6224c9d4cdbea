"""Seeded workload generators (ghcm_amd/csrc/field_gen.cpp) -- CPU."""
import hashlib
import json
import math
import os

import numpy as np

import ghcm_amd as g

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_ring_geometry():
    s = g.gen_ring_speed(256, 256, seed=7, wall_speed=1e-3)
    assert set(np.unique(s)) <= {1e-3, 1.0}
    assert s[128, 128] == 1.0                        # source cell is open
    yy, xx = np.mgrid[0:256, 0:256]
    r = np.hypot(xx - 128, yy - 128)
    for k in range(7):
        rk = 16 + k * (120 - 16) / 6
        ring = np.abs(r - rk) < 1.5
        frac_wall = (s[ring] == 1e-3).mean()
        assert 0.85 < frac_wall < 1.0                 # wall with one gap
    assert np.all(s[(r > 16 + 1.5) & (r < 33.33 - 1.5)] == 1.0)


def test_smooth_range_and_determinism():
    a = g.gen_smooth_speed(128, 96, seed=3, sigma=4.0)
    b = g.gen_smooth_speed(128, 96, seed=3, sigma=4.0)
    assert np.array_equal(a, b)
    assert a.min() == 0.1 and math.isclose(a.max(), 10.0, rel_tol=1e-12)
    la = np.log10(a)
    # smooth: neighbouring cells differ far less than the full span
    assert np.abs(np.diff(la, axis=1)).max() < 0.5
    c = g.gen_smooth_speed(128, 96, seed=4, sigma=4.0)
    assert not np.array_equal(a, c)


def test_points_in_grid():
    p = g.gen_points(100, 37, seed=11, n=500)
    assert p.shape == (500, 2)
    assert p[:, 0].max() < 100 and p[:, 1].max() < 37
    assert np.array_equal(p, g.gen_points(100, 37, seed=11, n=500))


def test_generator_fingerprints():
    """The workloads must not drift: digests of fixed-seed fields are committed."""
    fp = json.load(open(os.path.join(GOLDEN, "generator_fingerprints.json")))
    for name, spec in fp.items():
        if spec["kind"] == "ring":
            f = g.gen_ring_speed(spec["w"], spec["h"], spec["seed"], spec["wall_speed"], spec["scale"])
            digest = hashlib.sha256(f.tobytes()).hexdigest()
            assert digest == spec["sha256"], name
        elif spec["kind"] == "smooth":
            f = g.gen_smooth_speed(spec["w"], spec["h"], spec["seed"], spec["sigma"])
            # libm may differ in the last ulp across machines: compare rounded values
            digest = hashlib.sha256(np.round(np.log10(f), 9).tobytes()).hexdigest()
            assert digest == spec["sha256"], name
        else:
            p = g.gen_points(spec["w"], spec["h"], spec["seed"], spec["n"])
            assert p.tolist() == spec["points"], name
