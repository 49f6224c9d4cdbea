"""The batched (512 uneven instances) and served (Poisson arrivals, up to 4096^2) workloads at
their full sizes, every result against the oracle -- needs an MI355X.

Bar: each instance's times within 1e-4 * max T of the oracle's FP32 FMM (restating
src/eikonal_fmm.cpp:12-71), with the same finite/INF pattern, as for the plain solve.  The
oracle runs on the host's cores in threads (ctypes releases the GIL)."""
import concurrent.futures as cf
import os
import threading
import time

import numpy as np
import pytest

import ghcm_amd as g
import oracle
from ghcm_amd import workloads

pytestmark = pytest.mark.gpu
TOL = 1e-4
BATCH_CFG = dict(tile_w=64, tile_h=64, kernel=g.KERNEL_SWEEP, n_active=65536, inner_iter_cap=1)


def _cores():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except OSError:
        return 1


def _oracle_all(insts):
    with cf.ThreadPoolExecutor(max_workers=_cores()) as ex:
        return list(ex.map(lambda i: oracle.fmm_f32(i.cost, i.seeds), insts))


def _check(label, t, ref, seeds):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin), f"{label}: finite/INF pattern differs"
    scale = float(ref[fin].max()) if fin.any() else 0.0
    err = float(np.max(np.abs(t[fin].astype(np.float64) - ref[fin]))) if fin.any() else 0.0
    assert err <= TOL * scale + 1e-6, f"{label}: max|dT| {err} > {TOL} * {scale}"
    for x, y in np.asarray(seeds).reshape(-1, 2):
        assert t[y, x] == 0.0, f"{label}: seed ({x},{y})"
    return err / scale if scale else 0.0


@pytest.fixture(scope="module")
def batch512():
    t0 = time.perf_counter()
    insts = workloads.batch()
    print(f"batch of {len(insts)} generated in {time.perf_counter() - t0:.1f} s", flush=True)
    t0 = time.perf_counter()
    refs = _oracle_all(insts)
    print(f"oracle FMM over the batch in {time.perf_counter() - t0:.1f} s", flush=True)
    return insts, refs


@pytest.mark.timeout(900)
@pytest.mark.parametrize("launch_mode", [0, 1, 2, 4])
def test_batch512_every_instance_vs_oracle(batch512, launch_mode):
    """The 512-instance batch (sides 64..2048 per axis, ring or smooth, 1-8 sources) in one
    device wave loop: per-wave launches (0), one persistent cooperative launch (1), HIP-graph
    replay (2) and boundary-free waves (4)."""
    insts, refs = batch512
    print(f"case batch512 launch_mode={launch_mode}", flush=True)
    with g.Context(0, launch_mode=launch_mode, **BATCH_CFG) as ctx:
        dev = [(ctx.to_device(i.cost), ctx.alloc((i.h, i.w), np.float32)) for i in insts]
        st = ctx.solve_batch(g.FIELD_COST_F32, [(dc.ptr, dt.ptr, i.w, i.h, i.seeds)
                                                for (dc, dt), i in zip(dev, insts)])
        assert st["converged"] == 1
        print(f"  waves={st['waves']} visits={st['tile_visits']} device "
              f"{st['duration_device_us'] / 1e3:.2f} ms", flush=True)
        worst = 0.0
        for k, ((dc, dt), i, ref) in enumerate(zip(dev, insts, refs)):
            worst = max(worst, _check(f"instance {k} {i.w}x{i.h} {i.kind}", dt.download(), ref, i.seeds))
            dc.free()
            dt.free()
    print(f"  worst linf/max T over 512 instances: {worst:.2e}", flush=True)


@pytest.fixture(scope="module")
def serving_mix():
    """40 requests of the serving workload (square sides log-uniform 128..4096, ring or
    smooth, 1-4 sources) plus two 4096^2 ones: the plain smooth field and a ring field."""
    _, reqs = workloads.serving_requests(40, 1.0)
    reqs.insert(5, workloads.plain_4096())
    reqs.insert(20, workloads._make(4096, 4096, "ring", 6061, 3, 6062))
    t0 = time.perf_counter()
    refs = _oracle_all(reqs)
    print(f"serving mix of {len(reqs)} ({sum(r.w * r.h for r in reqs) / 1e6:.0f} Mcells), "
          f"oracle in {time.perf_counter() - t0:.1f} s", flush=True)
    return reqs, refs


@pytest.mark.timeout(900)
@pytest.mark.parametrize("rate", [150.0, 600.0])
def test_serving_poisson_arrivals_vs_oracle(serving_mix, rate):
    """Requests arrive at Poisson times (a client thread) while earlier ones are on the
    device; every third one goes through host buffers, the rest are resident in HBM.  Each
    result must match the oracle; requests must have been in flight together."""
    reqs, refs = serving_mix
    n = len(reqs)
    gaps = np.random.default_rng(int(rate)).exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    print(f"case serving poisson rate={rate}/s n={n}", flush=True)
    cfg = dict(BATCH_CFG, waves_per_poll=4)
    with g.Server(0, **cfg) as srv:
        ctx = g.Context(0)
        dev = {}
        for i, r in enumerate(reqs):
            if i % 3:
                dev[i] = (ctx.to_device(r.cost), ctx.alloc((r.h, r.w), np.float32))
        tickets = [None] * n

        def client():
            start = time.perf_counter()
            for i, r in enumerate(reqs):
                d = arrivals[i] - (time.perf_counter() - start)
                if d > 0:
                    time.sleep(d)
                if i in dev:
                    tickets[i] = srv.submit_device(dev[i][0].ptr, r.w, r.h, r.seeds, dev[i][1].ptr)
                else:
                    tickets[i] = srv.submit(r.cost, r.seeds)

        th = threading.Thread(target=client)
        th.start()
        out = [None] * n
        for i in range(n):
            while tickets[i] is None:
                time.sleep(0.0005)
            t, info = srv.wait(tickets[i], 120.0)
            assert info["status"] == g.OK
            out[i] = dev[i][1].download() if i in dev else t
        th.join()
        st = srv.stats()
        for d in dev.values():
            d[0].free()
            d[1].free()
        ctx.close()
    assert st["completed"] == n
    assert st["peak_live"] > 1, "requests were never in flight together"
    worst = 0.0
    for i, (r, ref) in enumerate(zip(reqs, refs)):
        worst = max(worst, _check(f"request {i} {r.w}x{r.h} {r.kind}", out[i], ref, r.seeds))
    print(f"  peak_live={st['peak_live']} worst linf/max T={worst:.2e}", flush=True)
