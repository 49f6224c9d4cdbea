"""Parity of the HIP path (libghcm.so) with the oracle -- needs an MI355X.

Bar: max |dT| <= 1e-4 * max T against the oracle's FP32 FMM (restating
src/eikonal_fmm.cpp), identical finite/INF pattern, T(seed) = 0, and the discrete
Godunov residual small at every cell.  Every call goes through the C-ABI.
"""
import os

import numpy as np
import pytest

import ghcm_amd as g
import oracle
from conftest import fixed_point_gap, ulp32

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-4          # max |dT| <= TOL * max T (north-star tolerance for the FP32 plain form)

TILES = [(8, 8), (16, 16), (32, 8), (32, 32)]
KERNELS = [g.KERNEL_SYSTOLIC, g.KERNEL_FIM]
SWEEP_TILES = [(64, 64), (64, 32), (64, 16), (32, 32)]


def check(t, ref, seeds, cost=None, tol=TOL, label="", eps=1e-5):
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(t)), f"{label}: INF pattern differs"
    if fin.any():
        err = float(np.max(np.abs(t[fin].astype(np.float64) - ref[fin])))
        assert err <= tol * float(ref[fin].max()) + 1e-6, f"{label}: max|dT|={err}"
    for x, y in np.asarray(seeds).reshape(-1, 2):
        assert t[y, x] == 0.0, f"{label}: seed ({x},{y}) = {t[y, x]}"
    assert np.all(t[fin] >= 0)
    if cost is not None:
        # converged to the discrete fixed point within a few ulps plus epsilon
        gap = fixed_point_gap(t, cost, seeds)
        assert np.all(gap <= 4 * ulp32(t) + 2 * eps), f"{label}: fixed-point gap {gap.max()}"


@pytest.fixture(scope="module")
def rings256():
    d = np.load(os.path.join(GOLDEN, "ring256_fmm64.npz"))
    speed = g.gen_ring_speed(256, 256, 1, 1e-3, 1.0)
    cost = g.cost_from_speed(speed)
    return cost, d["seeds"], oracle.fmm_f32(cost, d["seeds"]), d["time"]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("sched", [g.SCHED_DEVICE_TOPN, g.SCHED_HOST_PQ])
def test_rings256_every_config(rings256, kernel, tile, sched):
    cost, seeds, ref32, ref64 = rings256
    print(f"case rings256 kernel={kernel} tile={tile} sched={sched}", flush=True)
    with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=kernel, scheduler=sched,
                   compute_unit_overload=4.0) as ctx:
        t, st = ctx.solve(cost, seeds)
    assert st["converged"] == 1 and st["waves"] > 0
    check(t, ref32, seeds, cost, label="vs fmm32")
    check(t, ref64.astype(np.float64), seeds, label="vs fmm64 fixture")


@pytest.mark.parametrize("tile", SWEEP_TILES)
@pytest.mark.parametrize("sched,launch_mode", [(g.SCHED_DEVICE_TOPN, 0), (g.SCHED_DEVICE_TOPN, 1),
                                               (g.SCHED_DEVICE_TOPN, 2), (g.SCHED_DEVICE_TOPN, 3),
                                               (g.SCHED_DEVICE_TOPN, 4), (g.SCHED_DEVICE_TOPN, 5),
                                               (g.SCHED_HOST_PQ, 0), (g.SCHED_HOST_PQ, 2)])
@pytest.mark.parametrize("cap", [0, 1, 2])
def test_rings256_sweep_kernel(rings256, tile, sched, launch_mode, cap):
    cost, seeds, ref32, ref64 = rings256
    print(f"case rings256 sweep tile={tile} sched={sched} launch_mode={launch_mode} cap={cap}", flush=True)
    with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=g.KERNEL_SWEEP, scheduler=sched,
                   launch_mode=launch_mode, inner_iter_cap=cap, n_active=64) as ctx:
        t, st = ctx.solve(cost, seeds)
    assert st["converged"] == 1
    check(t, ref32, seeds, cost, label="sweep vs fmm32")
    check(t, ref64.astype(np.float64), seeds, label="sweep vs fmm64 fixture")


@pytest.mark.parametrize("kernel,tile", [(g.KERNEL_SYSTOLIC, (8, 8)), (g.KERNEL_FIM, (16, 16)),
                                         (g.KERNEL_SWEEP, (32, 32))])
def test_host_queue_graph_waves_equal_enqueued_waves(rings256, kernel, tile):
    """The host priority queue with each wave replayed as one captured graph (job upload,
    tile launch over n slots, status read-back) pops the same jobs in the same order as
    with three enqueued operations per wave, so with the one-wave tile kernels (systolic,
    FIM) its times are identical; the sweep kernel's four waves interleave their LDS
    updates freely, so it is held to the oracle.  A second solve of another size
    re-captures, and a third of the first size replays again."""
    cost, seeds, ref32, _ = rings256
    small = np.ascontiguousarray(cost[:200, :136])
    sseeds = np.array([[17, 150]], np.uint32)
    print(f"case host-queue graph waves kernel={kernel} tile={tile}", flush=True)
    out = {}
    for mode in (0, 2):
        with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=kernel, scheduler=g.SCHED_HOST_PQ,
                       launch_mode=mode, n_active=48) as ctx:
            a, sa = ctx.solve(cost, seeds)
            b, _ = ctx.solve(small, sseeds)
            c, sc = ctx.solve(cost, seeds)
        assert sa["converged"] == 1 and sa["waves"] > 10
        out[mode] = (a, b, c, sa["waves"], sa["tile_visits"])
    if kernel != g.KERNEL_SWEEP:
        for k in range(3):
            assert np.array_equal(out[0][k], out[2][k])
        assert out[0][3:] == out[2][3:]
        assert np.array_equal(out[2][0], out[2][2])
    check(out[2][0], ref32, seeds, cost, label="host-queue graph vs fmm32")
    check(out[2][2], ref32, seeds, cost, label="host-queue graph, replayed again")
    check(out[2][1], oracle.fmm_f32(small, sseeds), sseeds, small, label="host-queue graph, second size")


@pytest.mark.parametrize("sweep", [False, True])
@pytest.mark.parametrize("n_active", [1, 3, 64, 100000])
def test_top_n_sizes(rings256, n_active, sweep):
    """n smaller than the active set exercises the device top-n selection; with the 64x64
    sweep kernel, n far above the CU count selects its higher-occupancy form."""
    cost, seeds, ref32, _ = rings256
    cfg = dict(tile_w=64, tile_h=64, kernel=g.KERNEL_SWEEP, inner_iter_cap=1) if sweep else {}
    with g.Context(0, n_active=n_active, **cfg) as ctx:
        t, st = ctx.solve(cost, seeds)
    check(t, ref32, seeds, cost, label=f"n={n_active} sweep={sweep}")


@pytest.mark.parametrize("tile", [(64, 64), (64, 32)])
@pytest.mark.parametrize("n_active", [1, 3, 16, 64])
def test_boundary_free_top_n(tile, n_active):
    """launch_mode 5: the tile launch makes the top-n cut itself (every work-group applies the
    boundary kernel's selection rule to the same records, and spare work-groups carry the
    unselected tiles).  On a 1024^2 smooth field whose lists outgrow n, the solve matches the
    oracle, visits about as many tiles as the boundary kernel's selection (launch_mode 0) and
    no more than the boundary-free waves without a cut (launch_mode 4), which solve every
    queued tile."""
    speed = g.gen_smooth_speed(1024, 1024, 31, 32.0, 0.1, 10.0)
    cost = g.cost_from_speed(speed)
    seeds = g.gen_points(1024, 1024, 32, 3)
    ref = oracle.fmm_f32(cost, seeds)
    print(f"case boundary-free top-n tile={tile} n={n_active}", flush=True)
    st = {}
    for mode in (0, 4, 5):
        with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=g.KERNEL_SWEEP, inner_iter_cap=1,
                       n_active=n_active, launch_mode=mode) as ctx:
            t, st[mode] = ctx.solve(cost, seeds)
        assert st[mode]["converged"] == 1
        check(t, ref, seeds, cost, label=f"launch_mode {mode}")
    v0, v4, v5 = (st[m]["tile_visits"] for m in (0, 4, 5))
    print(f"  visits: mode0 {v0} mode4 {v4} mode5 {v5}; waves {st[0]['waves']} {st[4]['waves']} {st[5]['waves']}",
          flush=True)
    assert v5 <= v4
    assert abs(v5 - v0) <= 0.1 * v0 + 8


def test_long_lists_carried_by_the_tile_launch():
    """A 2048^2 ring field in 8x8 tiles at n = 64: lists grow to thousands of tiles, so the
    boundaries hand the unselected tiles to the tile launch, which carries them to the next
    wave; nothing may be lost or queued twice."""
    from ghcm_amd import workloads
    r = workloads._make(2048, 2048, "ring", 77, 4, 78)
    ref = oracle.fmm_f32(r.cost, r.seeds)
    print("case long lists systolic 8x8 2048^2 n=64", flush=True)
    with g.Context(0, tile_w=8, tile_h=8, kernel=g.KERNEL_SYSTOLIC, n_active=64) as ctx:
        t, st = ctx.solve(r.cost, r.seeds)
    assert st["converged"] == 1
    check(t, ref, r.seeds, r.cost, label="long lists systolic")


@pytest.mark.parametrize("launch_mode", [0, 1, 3, 4])
def test_long_first_list_in_every_launch_mode(launch_mode):
    """Every 32x32 tile of a 2048^2 field is a walled chamber (cost +INF on its last row and
    column) with a source at its centre: the first boundary selects 64 of 4,096 queued tiles,
    and a tile it does not carry is never queued again, since no neighbour can reach it.  The
    persistent (1) and fused (3) forms start with that boundary and their tile loops take no
    handed-off tiles, so the boundary must carry the unselected ones itself."""
    w = h = 2048
    cost = g.cost_from_speed(g.gen_smooth_speed(w, h, 91, 32.0, 0.1, 10.0))
    cost[31::32, :] = np.inf
    cost[:, 31::32] = np.inf
    seeds = [(32 * i + 15, 32 * j + 15) for j in range(h // 32) for i in range(w // 32)]
    ref = oracle.fmm_f32(cost, seeds)
    print(f"case walled chambers launch_mode={launch_mode}", flush=True)
    with g.Context(0, tile_w=32, tile_h=32, kernel=g.KERNEL_SWEEP, n_active=64, inner_iter_cap=1,
                   launch_mode=launch_mode) as ctx:
        t, st = ctx.solve(cost, seeds)
    assert st["converged"] == 1
    check(t, ref, seeds, cost, label=f"walled chambers launch_mode={launch_mode}")


def test_long_lists_carried_in_a_sweep_batch():
    """The same with the sweep kernel: 16 ring instances of 1024^2 in 32x32 tiles, n = 64."""
    from ghcm_amd import workloads
    reqs = [workloads._make(1024, 1024, "ring", 300 + i, 1 + i % 4, 400 + i) for i in range(16)]
    print("case long lists sweep batch 16 x 1024^2 n=64", flush=True)
    with g.Context(0, tile_w=32, tile_h=32, kernel=g.KERNEL_SWEEP, n_active=64,
                   inner_iter_cap=1) as ctx:
        dev = [(ctx.to_device(r.cost), ctx.alloc((r.h, r.w), np.float32)) for r in reqs]
        st = ctx.solve_batch(g.FIELD_COST_F32, [(dc.ptr, dt.ptr, r.w, r.h, r.seeds)
                                                for (dc, dt), r in zip(dev, reqs)])
        assert st["converged"] == 1
        for (dc, dt), r in zip(dev, reqs):
            check(dt.download(), oracle.fmm_f32(r.cost, r.seeds), r.seeds, r.cost,
                  label="long lists sweep batch")
            dc.free()
            dt.free()


SHAPES = [(1, 1), (1, 37), (37, 1), (13, 29), (100, 70), (3, 250), (257, 129)]


@pytest.mark.parametrize("kernel", KERNELS + [g.KERNEL_SWEEP])
@pytest.mark.parametrize("shape", SHAPES)
def test_ragged_obstacles_multi_seed(rng, kernel, shape):
    h, w = shape
    cost = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    cost[rng.random((h, w)) < 0.12] = np.inf
    seeds = {(0, 0), (w - 1, h - 1), (int(rng.integers(0, w)), int(rng.integers(0, h)))}
    seeds = sorted(seeds)
    for x, y in seeds:
        cost[y, x] = 1.0
    ref = oracle.fmm_f32(cost, seeds)
    tiles = ((64, 16), (32, 32)) if kernel == g.KERNEL_SWEEP else ((8, 8), (32, 8))
    for tile in tiles:
        for lm in ((0, 1, 4) if kernel == g.KERNEL_SWEEP else (0,)):
            print(f"case ragged {shape} kernel={kernel} tile={tile} launch_mode={lm}", flush=True)
            with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=kernel, n_active=7,
                           launch_mode=lm) as ctx:
                t, _ = ctx.solve(cost, seeds)
            check(t, ref, seeds, cost, label=f"{shape} {tile} lm={lm}")


def test_adjacent_seeds_and_edges():
    cost = np.full((40, 50), 2.0, dtype=np.float32)
    seeds = [(7, 7), (8, 7), (7, 8), (0, 39), (49, 0), (16, 16)]      # adjacent + tile corners
    ref = oracle.fmm_f32(cost, seeds)
    with g.Context(0) as ctx:
        t, _ = ctx.solve(cost, seeds)
    check(t, ref, seeds, cost)


def test_unreachable_region_stays_inf():
    cost = np.ones((64, 64), dtype=np.float32)
    cost[:, 20] = np.inf
    cost[30, 40:] = np.inf
    seeds = [(3, 3)]
    with g.Context(0, tile_w=16, tile_h=16) as ctx:
        t, _ = ctx.solve(cost, seeds)
    assert np.all(np.isinf(t[:, 20:]))
    check(t, oracle.fmm_f32(cost, seeds), seeds)


def test_no_seeds_all_inf():
    cost = np.ones((20, 20), dtype=np.float32)
    with g.Context(0) as ctx:
        t, st = ctx.solve(cost, np.zeros((0, 2), np.uint32))
    assert np.all(np.isinf(t)) and st["converged"] == 1


def test_strip_prefix_sums():
    cost = np.random.default_rng(1).uniform(0.1, 3.0, (1, 1000)).astype(np.float32)
    with g.Context(0, tile_w=32, tile_h=8) as ctx:
        t, _ = ctx.solve(cost, [(0, 0)])
    exp = np.concatenate([[0.0], np.cumsum(cost[0, 1:].astype(np.float64))])
    assert np.allclose(t[0], exp, rtol=1e-5)


def test_epsilon_zero_exact_convergence(rings256):
    cost, seeds, ref32, _ = rings256
    for kernel in KERNELS + [g.KERNEL_SWEEP]:
        tile = (64, 64) if kernel == g.KERNEL_SWEEP else (8, 8)
        with g.Context(0, min_change=0.0, kernel=kernel, tile_w=tile[0], tile_h=tile[1]) as ctx:
            t, _ = ctx.solve(cost, seeds)
        check(t, ref32, seeds, cost, tol=1e-5, label=f"eps=0 kernel={kernel}", eps=0.0)


def test_metamorphic_configs_agree(rng):
    speed = g.gen_smooth_speed(300, 220, seed=5, sigma=8.0)
    cost = g.cost_from_speed(speed)
    seeds = g.gen_points(300, 220, 6, 3)
    outs = []
    for tile in TILES:
        for kernel in KERNELS:
            with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=kernel, n_active=50) as ctx:
                outs.append(ctx.solve(cost, seeds)[0])
    for tile in SWEEP_TILES:
        with g.Context(0, tile_w=tile[0], tile_h=tile[1], kernel=g.KERNEL_SWEEP, n_active=50,
                       inner_iter_cap=1) as ctx:
            outs.append(ctx.solve(cost, seeds)[0])
    ref = oracle.fmm_f32(cost, seeds)
    for t in outs:
        check(t, ref, seeds, cost)
        assert np.max(np.abs(t - outs[0])) <= TOL * ref.max()


def test_cxx_api_path(rings256):
    cost, seeds, ref32, _ = rings256
    for kernel in KERNELS:
        t, st = g.hcm_solve_cxx(cost, seeds, block_size=16, compute_unit_overload=2.0,
                                kernel=kernel)
        check(t, ref32, seeds, label=f"c++ api kernel={kernel}")
        assert st["tile_w"] == 16


def test_errors_reported():
    cost = np.ones((16, 16), dtype=np.float32)
    with g.Context(0) as ctx:
        with pytest.raises(g.GhcmError) as e:
            ctx.solve(cost, [(16, 0)])
        assert e.value.status == g.ERR_INVALID_ARG
        ctx.set_config(max_waves=1, waves_per_poll=1)
        big = np.ones((256, 256), dtype=np.float32)
        with pytest.raises(g.GhcmError) as e:
            ctx.solve(big, [(0, 0)])
        assert e.value.status == g.ERR_NOT_CONVERGED
        ctx.set_config(max_waves=0, waves_per_poll=0)
        t, _ = ctx.solve(cost, [(0, 0)])           # context still usable
        assert t[0, 0] == 0.0
        # Widths above GHCM_MAX_WIDTH (4,194,304) are refused before any buffer is touched.
        with pytest.raises(g.GhcmError) as e:
            ctx.solve_device(0x1000, g.FIELD_COST_F32, 4194305, 1, [(0, 0)], 0x1000)
        assert e.value.status == g.ERR_INVALID_ARG and "width" in str(e.value)


def test_stats_of_a_new_context_are_its_own(rings256):
    """A context's first solve reports its own counters, not a previous context's (the host
    record lives in pinned memory that the runtime may hand out again)."""
    cost, seeds, ref32, _ = rings256
    with g.Context(0, compute_unit_overload=4.0) as ctx:
        _, st_small = ctx.solve(cost, seeds)
    with g.Context(0, tile_w=64, tile_h=64, kernel=g.KERNEL_SWEEP, inner_iter_cap=1) as ctx:
        _, st_big = ctx.solve(cost, seeds)
    # 8x8 tiles: 1,024 tiles, thousands of visits; 64x64 tiles: 16 tiles
    assert st_small["tile_visits"] > 1000
    assert 0 < st_big["tile_visits"] < 1000, st_big


def test_context_reuse_after_smaller_and_larger():
    with g.Context(0, tile_w=16, tile_h=16) as ctx:
        for (h, w) in ((300, 200), (17, 33), (512, 700)):
            cost = np.ones((h, w), dtype=np.float32)
            t, _ = ctx.solve(cost, [(w // 2, h // 2)])
            check(t, oracle.fmm_f32(cost, [(w // 2, h // 2)]), [(w // 2, h // 2)])
