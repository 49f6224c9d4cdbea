"""Seeded benchmark workloads of the gHCM path (fields and sources).

Plain (4096^2)   smooth log-normal speed in [0.1, 10], sigma 32, 4 point sources.
Rings (256^2)    seven ring walls (speed 1e-3) with one gap each, centre source.
Contrast (4096^2) the ring geometry scaled x16, walls at speed 1e-4 on background 1.
Batch            `n` instances, W and H drawn independently from {64, 128, ..., 2048},
                 1-8 sources, ring (rings scaled to min(W, H)/256, walls 1e-3) or smooth
                 (sigma 32, as the 4096^2 field) chosen per instance; one seed for all.
Serving          square sizes log-uniform in [128, 4096] (rounded to 64), 1-4 sources,
                 ring or smooth (sigma 32); Poisson arrival times at a given rate.
Volume (256^3)   3-D smooth log-normal speed in [0.1, 10]: iid N(0, 1) from numpy's PCG64
                 with the given seed, a Gaussian filter of sigma 8 (scipy.ndimage, reflect
                 boundary, truncated at 4 sigma), mapped log-linearly like the 2-D field;
                 4 point sources.

The per-instance streams use splitmix64 (the library's generator) through
``gen_points``-style draws so the same seed gives the same batch on any host.
"""
from __future__ import annotations

import concurrent.futures as cf
import math
from dataclasses import dataclass

import numpy as np

from . import cost_from_speed, gen_points, gen_ring_speed, gen_smooth_speed

PLAIN_SEED, PLAIN_POINT_SEED = 20261202, 20261203


@dataclass
class Instance:
    w: int
    h: int
    kind: str              # "ring" | "smooth"
    speed: np.ndarray      # (h, w) float64, or None when only the cost was kept
    seeds: np.ndarray      # (k, 2) uint32
    cost_f32: np.ndarray = None

    @property
    def cost(self) -> np.ndarray:
        if self.cost_f32 is None:
            self.cost_f32 = cost_from_speed(self.speed)
        return self.cost_f32


def plain_4096():
    speed = gen_smooth_speed(4096, 4096, PLAIN_SEED, 32.0, 0.1, 10.0)
    return Instance(4096, 4096, "smooth", speed, gen_points(4096, 4096, PLAIN_POINT_SEED, 4))


def smooth_speed_3d(n: int, seed: int, sigma: float = 8.0, lo: float = 0.1,
                    hi: float = 10.0) -> np.ndarray:
    """(n, n, n) smooth log-normal speed field in [lo, hi] (float32)."""
    from scipy.ndimage import gaussian_filter
    z = np.random.default_rng(seed).standard_normal((n, n, n), dtype=np.float32)
    z = gaussian_filter(z, sigma, mode="reflect", truncate=4.0)
    zmin, zmax = float(z.min()), float(z.max())
    e = np.log10(lo) + (np.log10(hi) - np.log10(lo)) * (z - zmin) / (zmax - zmin)
    return np.power(np.float32(10.0), e.astype(np.float32)).astype(np.float32)


def volume_256(seed: int = 7):
    """The 3-D workload: 256^3 smooth speed, 4 sources; returns (speed, seeds (x, y, z))."""
    n = 256
    speed = smooth_speed_3d(n, seed)
    pts = np.random.default_rng(seed + 1).integers(0, n, (4, 3)).astype(np.uint32)
    return speed, pts


def rings_256(seed: int = 1):
    speed = gen_ring_speed(256, 256, seed, 1e-3, 1.0)
    return Instance(256, 256, "ring", speed, np.array([[128, 128]], dtype=np.uint32))


def contrast_4096(seed: int = 5):
    speed = gen_ring_speed(4096, 4096, seed, 1e-4, 16.0)
    return Instance(4096, 4096, "ring", speed, np.array([[2048, 2048]], dtype=np.uint32))


def _draws(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from the library's splitmix64 stream (via gen_points)."""
    p = gen_points(1 << 30, 1, seed, n)          # x = floor(u * 2^30)
    return p[:, 0].astype(np.float64) / float(1 << 30)


def _make(w: int, h: int, kind: str, fseed: int, n_src: int, pseed: int, sigma: float = 32.0,
          cost_only: bool = False) -> Instance:
    if kind == "ring":
        speed = gen_ring_speed(w, h, fseed, 1e-3, min(w, h) / 256.0)
    else:
        speed = gen_smooth_speed(w, h, fseed, sigma, 0.1, 10.0)
    inst = Instance(w, h, kind, speed, gen_points(w, h, pseed, n_src))
    if cost_only:
        inst.cost_f32 = cost_from_speed(speed)
        inst.speed = None
    return inst


def batch(n: int = 512, seed: int = 4040, max_side: int = 2048, workers: int = 8):
    """The uneven batch: sides from {64, ..., max_side}, 1-8 sources, ring or smooth
    (costs only: the speed fields are dropped once converted)."""
    u = _draws(seed, 5 * n).reshape(n, 5)
    sides = np.arange(64, max_side + 1, 64)
    specs = []
    for i in range(n):
        w = int(sides[min(int(u[i, 0] * len(sides)), len(sides) - 1)])
        h = int(sides[min(int(u[i, 1] * len(sides)), len(sides) - 1)])
        n_src = 1 + min(int(u[i, 2] * 8), 7)
        kind = "ring" if u[i, 3] < 0.5 else "smooth"
        specs.append((w, h, kind, seed * 1000 + 2 * i, n_src, seed * 1000 + 2 * i + 1, 32.0, True))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda s: _make(*s), specs))


def serving_requests(n: int, rate_per_s: float, seed: int = 5050, lo: int = 128, hi: int = 4096,
                     workers: int = 8):
    """(arrival_times, instances): square sides log-uniform in [lo, hi], Poisson arrivals."""
    u = _draws(seed, 4 * n).reshape(n, 4)
    specs = []
    for i in range(n):
        side = int(round(math.exp(math.log(lo) + u[i, 0] * (math.log(hi) - math.log(lo))) / 64.0)) * 64
        side = max(64, min(hi, side))
        n_src = 1 + min(int(u[i, 1] * 4), 3)
        kind = "ring" if u[i, 2] < 0.5 else "smooth"
        specs.append((side, side, kind, seed * 1000 + 2 * i, n_src, seed * 1000 + 2 * i + 1, 32.0, True))
    gaps = -np.log(1.0 - u[:, 3]) / rate_per_s
    arrivals = np.cumsum(gaps) - gaps[0]
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        inst = list(ex.map(lambda s: _make(*s), specs))
    return arrivals, inst
