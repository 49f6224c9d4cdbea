"""Three-dimensional solves (the K = 3 generalisation of src/eikonal.h:73-118) on the GPU,
checked against the oracle's 3-D FMM in FP64 -- needs an MI355X."""
import functools

import numpy as np
import pytest

import ghcm_amd as g
import oracle
from conftest import fixed_point_gap3d, ulp32

pytestmark = pytest.mark.gpu

TOL = 1e-4      # max |dT| <= TOL * max T, as for the plain 2-D form

# The two 3-D tile solvers: 8x8x8 tiles relaxed by in-tile passes (every kernel but the
# sweep), and 64x8x8 tiles relaxed by the eight octant sweeps (the context's 2-D tile shape
# must still be one the sweep kernel is built for), the latter relaxed until converged in
# each visit (cap 0), one round per visit (cap 1: the change test made in the commit pass,
# the configuration the volume is measured in) and two (cap 2: a round check, then a capped
# last round).
KERNELS = {"passes": dict(kernel=g.KERNEL_SYSTOLIC),
           "sweep": dict(kernel=g.KERNEL_SWEEP, tile_w=64, tile_h=64),
           "sweep_cap1": dict(kernel=g.KERNEL_SWEEP, tile_w=64, tile_h=64, inner_iter_cap=1),
           "sweep_cap2": dict(kernel=g.KERNEL_SWEEP, tile_w=64, tile_h=64, inner_iter_cap=2)}
SHAPES = {"passes": (8, 8, 8), "sweep": (64, 8, 8), "sweep_cap1": (64, 8, 8), "sweep_cap2": (64, 8, 8)}


def _check(t, ref, cost, seeds, eps):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin)
    for x, y, z in np.asarray(seeds).reshape(-1, 3):
        assert t[z, y, x] == 0.0
    err = np.max(np.abs(t[fin].astype(np.float64) - ref[fin])) if fin.any() else 0.0
    assert err <= TOL * ref[fin].max(), err
    gap = fixed_point_gap3d(t, cost, seeds)
    assert np.all(gap <= 4 * ulp32(t) + 2 * eps + 1e-6)
    return err


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("shape", [(8, 8, 8), (17, 23, 29), (40, 9, 33), (9, 17, 130)])
def test_3d_parity_random_with_walls(shape, kernel, rng):
    d, h, w = shape
    print(f"case 3d random {shape} {kernel}", flush=True)
    cost = rng.uniform(0.1, 10.0, shape).astype(np.float32)
    if d > 8:
        cost[d // 2, 1:h - 1, 2:w - 2] = np.inf          # a wall with a rim to go around
    seeds = [(0, 0, 0), (w - 1, h - 1, d - 1), (w // 2, 0, d // 3)]
    ref = oracle.fmm3d_f64(cost, seeds)
    with g.Context(0, n_active=64, **KERNELS[kernel]) as ctx:
        t, st = ctx.solve3d(cost, seeds)
    assert st["converged"] == 1
    assert (st["tile_w"], st["tile_h"], st["tile_d"]) == SHAPES[kernel]
    _check(t, ref, cost, seeds, ctx.config.min_change)


def test_3d_seeds_on_tile_edge_columns(rng):
    """Seeds on the first and last x-column of 64x8x8 sweep tiles: the seed sets the tile's
    edge column too, which its x-neighbours stage as their halo; a ragged last tile (x 192..199)
    has no last column of its own to publish."""
    cost = rng.uniform(0.2, 4.0, (12, 20, 200)).astype(np.float32)
    seeds = [(63, 3, 2), (64, 17, 9), (127, 0, 0), (128, 19, 11), (192, 5, 1), (199, 10, 5)]
    ref = oracle.fmm3d_f64(cost, seeds)
    print("case 3d seeds on edge columns", flush=True)
    with g.Context(0, n_active=64, **KERNELS["sweep"]) as ctx:
        t, st = ctx.solve3d(cost, seeds)
    assert st["converged"] == 1
    _check(t, ref, cost, seeds, ctx.config.min_change)


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_3d_slab_matches_2d(kernel, rng):
    cost2 = rng.uniform(0.1, 10.0, (64, 80)).astype(np.float32)
    seeds2 = [(3, 4), (70, 50)]
    print(f"case 3d slab {kernel}", flush=True)
    with g.Context(0) as ctx:
        t2, _ = ctx.solve(cost2, seeds2)
    with g.Context(0, **KERNELS[kernel]) as ctx:
        t3, _ = ctx.solve3d(cost2[None], [(x, y, 0) for x, y in seeds2])
    ref = oracle.fmm_f64(cost2, seeds2)
    assert np.max(np.abs(t3[0].astype(np.float64) - ref)) <= TOL * ref.max()
    assert np.max(np.abs(t3[0].astype(np.float64) - t2)) <= 1e-5 * ref.max()


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("kind", ["speed_f32", "speed_f16", "speed_bf16"])
def test_3d_speed_kinds(kind, kernel):
    from ghcm_amd import workloads
    speed = workloads.smooth_speed_3d(48, 3, sigma=4.0)
    seeds = [(5, 40, 20), (30, 2, 44)]
    if kind == "speed_f16":
        stored = speed.astype(np.float16)
        field = stored
        eff = stored.astype(np.float32)
    elif kind == "speed_bf16":
        field = g.to_bf16(speed)
        eff = (field.astype(np.uint32) << 16).view(np.float32)
    else:
        field = eff = speed
    cost = (1.0 / eff.astype(np.float64))
    ref = oracle.fmm3d_f64(cost, seeds)
    print(f"case 3d kind {kind} {kernel}", flush=True)
    with g.Context(0, **KERNELS[kernel]) as ctx:
        t, _ = ctx.solve3d(field, seeds, kind=g.KINDS[kind])
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin)
    assert np.max(np.abs(t[fin] - ref[fin])) <= TOL * ref[fin].max()


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_3d_graph_mode_and_device_entry(kernel, rng):
    cost = rng.uniform(0.5, 2.0, (24, 40, 140)).astype(np.float32)
    seeds = [(1, 2, 3), (138, 30, 20)]
    ref = oracle.fmm3d_f64(cost, seeds)
    print(f"case 3d graph + device {kernel}", flush=True)
    with g.Context(0, launch_mode=2, waves_per_poll=4, **KERNELS[kernel]) as ctx:
        t_graph, st = ctx.solve3d(cost, seeds)
        assert st["converged"] == 1
        dc = ctx.to_device(cost)
        dt = ctx.alloc(cost.shape, np.float32)
        ctx.solve3d_device(dc.ptr, g.FIELD_COST_F32, 140, 40, 24, seeds, dt.ptr)
        t_dev = dt.download()
    for t in (t_graph, t_dev):
        assert np.max(np.abs(t - ref)) <= TOL * ref.max()


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_3d_errors(kernel):
    cost = np.ones((4, 4, 4), np.float32)
    with g.Context(0, scheduler=g.SCHED_HOST_PQ, **KERNELS[kernel]) as ctx:
        with pytest.raises(g.GhcmError) as e:
            ctx.solve3d(cost, [(0, 0, 0)])
        assert e.value.status == g.ERR_UNSUPPORTED
    with g.Context(0, **KERNELS[kernel]) as ctx:
        with pytest.raises(g.GhcmError) as e:
            ctx.solve3d(cost, [(0, 0, 4)])
        assert e.value.status == g.ERR_INVALID_ARG


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_3d_volume_128_vs_fp64(kernel):
    from ghcm_amd import workloads
    speed = workloads.smooth_speed_3d(128, 11)
    cost = (1.0 / speed).astype(np.float32)
    seeds = [(10, 20, 30), (100, 110, 90), (64, 64, 64)]
    ref = oracle.fmm3d_f64(cost.astype(np.float64), seeds)
    print(f"case 3d volume 128 {kernel}", flush=True)
    with g.Context(0, **KERNELS[kernel]) as ctx:
        t, st = ctx.solve3d(cost, seeds)
    assert st["converged"] == 1
    _check(t, ref, cost, seeds, ctx.config.min_change)


@functools.lru_cache(maxsize=1)
def _volume_256():
    from ghcm_amd import workloads
    speed, seeds = workloads.volume_256()
    cost = (np.float32(1.0) / speed).astype(np.float32)
    return cost, seeds, oracle.fmm3d_f64(cost.astype(np.float64), seeds)


@pytest.mark.parametrize("kernel,n_active", [("passes", 4096), ("sweep", 256), ("sweep_cap1", 2048)])
def test_3d_long_lists_carried_by_the_tile_launch(kernel, n_active):
    """The 256^3 bench volume: with 8x8x8 tiles at n = 4096 its boundaries select from lists
    of 4-6 thousand tiles, with 64x8x8 tiles at n = 256 from lists of several hundred, and
    hand the unselected ones to the tile launch, which carries them to the next wave;
    nothing may be lost or queued twice.  The last case is the measured volume configuration
    (one round per visit, n = 2048)."""
    cost, seeds, ref = _volume_256()
    print(f"case 3d long lists 256^3 {kernel} n={n_active}", flush=True)
    with g.Context(0, n_active=n_active, **KERNELS[kernel]) as ctx:
        t, st = ctx.solve3d(cost, seeds)
    assert st["converged"] == 1
    _check(t, ref, cost, seeds, ctx.config.min_change)
