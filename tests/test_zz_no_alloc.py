"""The allocation-free entry point (ghcm_solve_device_noalloc) against ghcm_solve_device --
needs an MI355X.  Both run the same kernels; the allocation-free one keeps the scheduler
state in the device's static scratch array.  Two calls in a row, the second on a smaller
grid (the array keeps the first call's contents), each checked against the allocating entry
point and the oracle within 1e-4 * max T with the same finite/INF pattern."""
import numpy as np
import pytest

import ghcm_amd as g
import oracle
from ghcm_amd import workloads

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _solve_both(ctx, cost, seeds):
    h, w = cost.shape
    dc = ctx.to_device(cost)
    d1 = ctx.alloc((h, w), np.float32)
    d2 = ctx.alloc((h, w), np.float32)
    st1 = ctx.solve_device(dc.ptr, g.FIELD_COST_F32, w, h, seeds, d1.ptr)
    st2 = ctx.solve_device_noalloc(dc.ptr, g.FIELD_COST_F32, w, h, seeds, d2.ptr)
    t1, t2 = d1.download(), d2.download()
    for d in (dc, d1, d2):
        d.free()
    return t1, t2, st1, st2


def _close(t, ref, label):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin), f"{label}: finite/INF pattern differs"
    scale = float(ref[fin].max())
    err = float(np.max(np.abs(t[fin].astype(np.float64) - ref[fin].astype(np.float64))))
    assert err <= TOL * scale, f"{label}: max|dT| {err} > {TOL} * {scale}"


@pytest.mark.parametrize("launch_mode", [0, 4])
def test_noalloc_matches_solve_device_twice(launch_mode):
    big = workloads._make(1536, 1024, "smooth", 811, 3, 812)
    small = workloads._make(300, 200, "ring", 813, 1, 814)
    with g.Context(0, tile_w=64, tile_h=64, kernel=g.KERNEL_SWEEP, n_active=512,
                   inner_iter_cap=1, launch_mode=launch_mode) as ctx:
        for label, r in (("first call 1536x1024", big), ("second call 300x200", small)):
            print(f"case noalloc launch_mode={launch_mode} {label}", flush=True)
            t1, t2, st1, st2 = _solve_both(ctx, r.cost, r.seeds)
            assert st1["converged"] == 1 and st2["converged"] == 1
            assert st2["tiles_total"] == st1["tiles_total"] and st2["waves"] > 0
            _close(t2, t1, f"{label}: noalloc vs solve_device")
            _close(t2, oracle.fmm_f32(r.cost, r.seeds), f"{label}: noalloc vs oracle")
            for x, y in r.seeds:
                assert t2[y, x] == 0.0


def test_noalloc_reference_default_tiles_and_limits():
    """The reference-default 8x8 systolic tiles, then the refusals: a grid whose tiles do not
    fit the scratch array, and a configuration that would allocate (graph replay)."""
    r = workloads._make(640, 480, "ring", 815, 2, 816)
    with g.Context(0, compute_unit_overload=4.0) as ctx:
        print("case noalloc 8x8 systolic 640x480", flush=True)
        t1, t2, _, st2 = _solve_both(ctx, r.cost, r.seeds)
        assert st2["converged"] == 1
        _close(t2, t1, "8x8: noalloc vs solve_device")
        # 8192^2 in 8x8 tiles: 1,048,576 tiles, far above the array; refused before any
        # device work (the pointers are never touched)
        with pytest.raises(g.GhcmError) as e:
            ctx.solve_device_noalloc(0x1000, g.FIELD_COST_F32, 8192, 8192, [(0, 0)], 0x1000)
        assert e.value.status == g.ERR_INVALID_ARG and "scratch" in str(e.value)
        ctx.set_config(launch_mode=2)
        with pytest.raises(g.GhcmError) as e:
            ctx.solve_device_noalloc(0x1000, g.FIELD_COST_F32, 64, 64, [(0, 0)], 0x1000)
        assert e.value.status == g.ERR_UNSUPPORTED
