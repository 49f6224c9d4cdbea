#!/usr/bin/env python3
"""Benchmark of the gHCM hot path on one MI355X.

Workload (the 4096^2 plain solve): a 4096 x 4096 FP32 cost field, cost = 1/speed with
speed = 10^(-1 + 2 (z - min z)/(max z - min z)), z a standard-normal field (seed
20261202) smoothed by a Gaussian of sigma 32 cells; 4 point sources drawn uniformly
(seed 20261203).  A "step" is one complete solve: initialise the time field, run the
device wave loop until the active-tile list drains.  Field and time buffers are resident
in HBM before the timed region starts.

Prints ONE JSON line on stdout: Mcells/s of the whole solve, the roofline entry of the
dominant kernel (the tile-solve kernel: algorithmic bytes per launch / average launch
duration, from HIP events stamped by each dispatch on the solver stream), and the CPU
baseline (the oracle's FP32 FMM on this host, one core; the FP32 FIM beside it).  Stamping
every launch perturbs the loop, so `value` comes from the un-instrumented timed steps and
the roofline from the same number of instrumented steps run right after them.

Beside the headline the line carries:
  batched    512 uneven instances in one device wave loop (autotuned over tile shapes,
             n, round caps and launch modes) vs the same solves launched one at a time,
             the L_inf of every batched and one-at-a-time result vs the oracle's FP64 FMM,
             the batch's algorithmic bytes over its wall time (fraction of HBM peak), and
             the oracle FMM on all cores;
  serving    the continuous-batching server on 2,048 requests resident in HBM and 256 on
             host buffers: saturated solves/s vs one-at-a-time, p50/p99 latency under
             Poisson arrivals at 50 % and 80 % of the saturated rate, served results vs
             the FP64 FMM, and the oracle FMM over the same fields on all cores;
  precision  the 4096^2 smooth field and a 4096^2 ring field of contrast 1e4, each stored
             as FP32 cost and FP32/FP16/BF16 speed, L_inf/L1 vs the FP64 FMM and the FP32
             FIM;
  reference_kernels  the 4096^2 solve with the systolic (8x8, 32x8) and FIM (16x16) tile
             kernels at the reference's default wave size, each with its roofline;
  launch_modes  the same 4096^2 solve and configuration in the boundary-free wave loops:
             launch_mode 4 (every queued tile each wave) and 5 (the top-n cut made by the
             tile launch), time per solve, waves, visits and per-launch time;
  volume     the 256^3 three-dimensional solve with the octant-sweep kernel (64x8x8
             tiles), its roofline and parity vs the FP64 3-D FMM, the in-tile-pass kernel
             (8x8x8 tiles) timed beside it, and the oracle's FP32 3-D FMM on one core.
Every CPU baseline names the host CPU model and the cores it used.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
W = H = 4096
FIELD_SEED, POINT_SEED = 20261202, 20261203
METRIC = "Mcells/s at 4096² (% HBM peak); batched solves/s vs plain; p99 latency"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_solve_sweep_plain4096_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per k_solve launch from the committed rocprofv3 --pmc summary, if any."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def host_description() -> dict:
    """The CPU the baselines run on: model name, logical CPUs (nproc) and the CPUs this
    process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except OSError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "cpus_allowed": allowed}


def _with_host(d: dict) -> dict:
    d.update({k: v for k, v in host_description().items() if k in ("cpu_model", "nproc")})
    return d


def cpu_baseline(cost, seeds, trials: int):
    """The oracle's FP32 FMM (restating src/eikonal_fmm.cpp) on one host core."""
    import oracle
    old = os.sched_getaffinity(0)
    try:
        os.sched_setaffinity(0, {min(old)})
    except OSError:
        pass
    times, ref = [], None
    try:
        for _ in range(trials):
            t0 = time.perf_counter()
            ref = oracle.fmm_f32(cost, seeds)
            times.append(time.perf_counter() - t0)
    finally:
        try:
            os.sched_setaffinity(0, old)
        except OSError:
            pass
    med = statistics.median(times)
    return ref, _with_host({"value": round(W * H / med / 1e6, 3), "unit": "Mcells/s", "cores": 1,
                            "kind": "port",
                            "sample": f"full {W}x{H} workload, oracle FP32 FMM, {trials} trials, "
                                      f"median {med:.2f} s, pinned to one core"})


BATCH_CANDIDATES = [  # autotuner search space for the batch (kept small so the bench stays short)
    dict(tile_w=64, tile_h=64, kernel=2, n_active=16384, inner_iter_cap=1, launch_mode=0),
    dict(tile_w=64, tile_h=64, kernel=2, n_active=65536, inner_iter_cap=1, launch_mode=0),
    dict(tile_w=64, tile_h=64, kernel=2, n_active=65536, inner_iter_cap=1, launch_mode=2),
    dict(tile_w=64, tile_h=64, kernel=2, n_active=65536, inner_iter_cap=2, launch_mode=2),
    dict(tile_w=64, tile_h=64, kernel=2, n_active=2048, inner_iter_cap=1, launch_mode=1),
    dict(tile_w=64, tile_h=64, kernel=2, n_active=65536, inner_iter_cap=1, launch_mode=4),
    dict(tile_w=64, tile_h=32, kernel=2, n_active=65536, inner_iter_cap=1, launch_mode=2),
    dict(tile_w=32, tile_h=32, kernel=2, n_active=65536, inner_iter_cap=1, launch_mode=2),
    dict(tile_w=16, tile_h=16, kernel=0, n_active=65536, inner_iter_cap=0, launch_mode=2),
    dict(tile_w=32, tile_h=8, kernel=0, n_active=65536, inner_iter_cap=0, launch_mode=2),
    dict(tile_w=8, tile_h=8, kernel=0, n_active=65536, inner_iter_cap=0, launch_mode=2),
]


def host_cores() -> int:
    """Cores this process may use (the GPU box grants 16 per GPU)."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except OSError:
        return 1


def _oracle_map(fn, items):
    """fn over items on the host's cores (ctypes releases the GIL)."""
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max_workers=host_cores()) as ex:
        return list(ex.map(fn, items))


def accuracy_vs_fp64(results, refs64) -> dict:
    """L_inf / max T and mean L1 of each result against its FP64 FMM, worst and mean over
    the results (finite cells; a finite/INF mismatch counts)."""
    import oracle
    rel, l1, mism = [], [], 0
    for t, r in zip(results, refs64):
        m = oracle.error_metrics(t, r)
        rel.append(m["linf"] / max(m["max_ref"], 1e-30))
        l1.append(m["l1"])
        mism += m["finiteness_mismatch"]
    return {"linf_rel_vs_fp64": float(max(rel)), "linf_rel_vs_fp64_mean": float(sum(rel) / len(rel)),
            "l1_vs_fp64_mean": float(sum(l1) / len(l1)), "finiteness_mismatch": mism,
            "results": len(rel)}


def bench_batch(g, plain_cfg: dict, steps: int, autotune: bool = True) -> dict:
    """512 uneven instances in one device wave loop vs the same solves one at a time.
    autotune=False skips the search and solves with BATCH_CANDIDATES[1], the configuration
    the search picks on MI355X (for counter runs that should see only the batch's kernels)."""
    import numpy as np
    from ghcm_amd import workloads
    t0 = time.perf_counter()
    insts = workloads.batch()
    log(f"bench: batch of {len(insts)} generated in {time.perf_counter() - t0:.1f} s")
    cells = sum(i.w * i.h for i in insts)
    with g.Context(0, **plain_cfg) as ctx:
        dev = [(ctx.to_device(i.cost), ctx.alloc((i.h, i.w), np.float32)) for i in insts]
        args = [(dc.ptr, dt.ptr, i.w, i.h, i.seeds) for (dc, dt), i in zip(dev, insts)]
        for a in args[:8]:                                   # warm-up
            ctx.solve_device(a[0], g.FIELD_COST_F32, a[2], a[3], a[4], a[1])
        t0 = time.perf_counter()
        for a in args:
            ctx.solve_device(a[0], g.FIELD_COST_F32, a[2], a[3], a[4], a[1])
        t_plain = time.perf_counter() - t0
        ref = [dt.download() for _, dt in dev]          # the one-at-a-time results
        if autotune:
            best, res = ctx.autotune(g.FIELD_COST_F32, args, BATCH_CANDIDATES, reps=1)
            log(f"bench: batch autotune -> {best}")
        else:
            best, res = dict(BATCH_CANDIDATES[1]), []
            ctx.set_config(**best)
            log(f"bench: batch configuration (no search) -> {best}")
        prepared = ctx.prepare_batch(args)          # the instance array, built once
        ctx.solve_batch(g.FIELD_COST_F32, prepared)
        ts = []
        for _ in range(max(1, steps)):
            t0 = time.perf_counter()
            st = ctx.solve_batch(g.FIELD_COST_F32, prepared)
            ts.append(time.perf_counter() - t0)
        t_batch = statistics.median(ts)
        worst = 0.0
        batched = []
        for (_, dt), r in zip(dev, ref):
            t = dt.download()
            batched.append(t)
            fin = np.isfinite(r)
            if not np.array_equal(fin, np.isfinite(t)):
                worst = float("inf")
            elif fin.any():
                worst = max(worst, float(np.max(np.abs(t[fin] - r[fin])) / max(1e-30, r[fin].max())))
        for dc, dt in dev:
            dc.free()
            dt.free()
    # CPU: the oracle FMM over independent instances on every host core, on a sample.
    import oracle
    cores = host_cores()
    sample = insts[:64]
    t0 = time.perf_counter()
    _oracle_map(lambda i: oracle.fmm_f32(i.cost, i.seeds), sample)
    t_cpu = time.perf_counter() - t0
    # Accuracy of every batched result and every one-at-a-time result against the FP64 FMM
    # (north_star: batched solves/s at equal accuracy against the FP64 oracle).
    t0 = time.perf_counter()
    refs64 = _oracle_map(lambda i: oracle.fmm_f64(i.cost.astype(np.float64), i.seeds), insts)
    acc = {"batched": accuracy_vs_fp64(batched, refs64), "plain": accuracy_vs_fp64(ref, refs64),
           "reference": f"oracle FP64 FMM of every instance ({time.perf_counter() - t0:.1f} s on "
                        f"{cores} threads)"}
    del refs64, batched, ref
    keys = ("tile_w", "tile_h", "kernel", "n_active", "inner_iter_cap", "launch_mode")
    return {"workload": f"{len(insts)} uneven instances, sides 64..2048, {cells / 1e6:.0f} Mcells, "
                        "ring or smooth, 1-8 sources (seed 4040)",
            "cpu_baseline": _with_host({"value": round(len(sample) / t_cpu, 2), "unit": "solves/s",
                                        "cores": cores, "kind": "port",
                                        "sample": f"oracle FP32 FMM over the first {len(sample)} "
                                                  f"batch instances, {cores} threads, {t_cpu:.1f} s"}),
            "accuracy": acc,
            "linf_rel_vs_fp64": acc["batched"]["linf_rel_vs_fp64"],
            "solves_per_s": round(len(insts) / t_batch, 1),
            "plain_solves_per_s": round(len(insts) / t_plain, 1),
            "speedup_vs_plain": round(t_plain / t_batch, 2),
            "batch_ms": round(t_batch * 1e3, 3), "plain_ms": round(t_plain * 1e3, 3),
            "batch_mcells_per_s": round(cells / t_batch / 1e6, 1),
            "waves": st["waves"], "tile_visits": st["tile_visits"], "converged": st["converged"],
            # Algorithmic bytes of the whole batched solve (visits x bytes per visit) over its
            # wall time: the batch's average rate, boundary launches and gaps included.
            "algorithmic_gb_per_s": round(st["tile_visits"] * st["bytes_per_visit"] / t_batch / 1e9, 1),
            "frac_of_hbm_peak": round(st["tile_visits"] * st["bytes_per_visit"] / t_batch / 1e9
                                      / HBM_PEAK_GBS, 4),
            "autotuned": {k: best[k] for k in keys},
            "autotune_device_ms": {f"{c['tile_w']}x{c['tile_h']}/k{c['kernel']}/n{c['n_active']}/"
                                   f"cap{c['inner_iter_cap']}/lm{c['launch_mode']}":
                                   round(c["device_us"] / 1e3, 3) for c in res},
            "max_rel_diff_vs_plain": worst}


def _serve_runs(srv, submit, n_requests: int, seed: int, loads=(0.8,)) -> dict:
    """Saturated run (every request submitted at once), then Poisson arrivals at each of
    `loads` times the saturated rate; latencies are submit -> result, measured by the library."""
    import threading

    import numpy as np
    for _ in range(2):          # the first pass grows the server's pinned staging pool
        t0 = time.perf_counter()
        tickets = [submit(i) for i in range(n_requests)]
        results = [srv.wait(t) for t in tickets]
        t_sat = time.perf_counter() - t0
    sat_rate = n_requests / t_sat
    out = {"results": results, "saturated_solves_per_s": round(sat_rate, 1), "requests": n_requests,
           "loads": {}}
    for li, load in enumerate(loads):
        rate = load * sat_rate
        gaps = np.random.default_rng(seed + li).exponential(1.0 / rate, n_requests)
        arrivals = np.cumsum(gaps) - gaps[0]
        tickets = [None] * n_requests

        def client():
            start = time.perf_counter()
            for i in range(n_requests):
                d = arrivals[i] - (time.perf_counter() - start)
                if d > 0:
                    time.sleep(d)
                tickets[i] = submit(i)

        th = threading.Thread(target=client)
        t0 = time.perf_counter()
        th.start()
        lat = []
        for i in range(n_requests):
            while tickets[i] is None:
                time.sleep(0.0002)
            _, info = srv.wait(tickets[i])
            lat.append(info["latency_us"] / 1e3)
        th.join()
        t_open = time.perf_counter() - t0
        lat = np.array(lat)
        out["loads"][f"{load:.2f}"] = {
            "poisson_rate_per_s": round(rate, 1),
            "poisson_achieved_per_s": round(n_requests / t_open, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3)}
    out["peak_live"] = srv.stats()["peak_live"]
    return out


def bench_serving(g, serve_cfg: dict, n_distinct: int = 128, n_requests: int = 2048,
                  n_host_requests: int = 256) -> dict:
    """Continuous batching on requests already in HBM (device buffers) and on host buffers
    (PCIe in and out), each against the same solves made one at a time; every served result
    of the distinct fields is checked against the FP64 FMM, and the oracle FMM on all host
    cores is the CPU leg."""
    import numpy as np

    import oracle
    from ghcm_amd import workloads
    t0 = time.perf_counter()
    _, insts = workloads.serving_requests(n_distinct, 1.0)
    log(f"bench: {n_distinct} serving fields generated in {time.perf_counter() - t0:.1f} s")
    pick = [i % n_distinct for i in range(n_requests)]
    loads = (0.5, 0.8)

    def rel_diff(t, p):
        fin = np.isfinite(p)
        if not np.array_equal(fin, np.isfinite(t)):
            return float("inf")
        return float(np.max(np.abs(t[fin] - p[fin])) / p[fin].max()) if fin.any() else 0.0

    cells = sum(r.w * r.h for r in insts)
    out = {"workload": f"{n_requests} requests in HBM ({n_host_requests} more with host buffers) "
                       f"cycling {n_distinct} square fields ({cells / 1e6:.0f} Mcells), sides "
                       "log-uniform 128..4096, ring or smooth (sigma 32), 1-4 sources (seed "
                       "5050); Poisson arrivals at 50% and 80% of the saturated rate"}
    with g.Context(0, **serve_cfg) as ctx:
        # ---- device-resident requests ----
        d_fields = [ctx.to_device(r.cost) for r in insts]
        d_plain = [ctx.alloc((r.h, r.w), np.float32) for r in insts]
        d_times = [ctx.alloc((insts[k].h, insts[k].w), np.float32) for k in pick]
        for k in range(4):
            r = insts[k]
            ctx.solve_device(d_fields[k].ptr, g.FIELD_COST_F32, r.w, r.h, r.seeds, d_plain[k].ptr)
        t0 = time.perf_counter()
        for k, r in enumerate(insts):
            ctx.solve_device(d_fields[k].ptr, g.FIELD_COST_F32, r.w, r.h, r.seeds, d_plain[k].ptr)
        t_plain_dev = (time.perf_counter() - t0) / n_distinct
        plain = [d.download() for d in d_plain]
        with g.Server(0, **serve_cfg) as srv:
            def submit_dev(i):
                k = pick[i]
                r = insts[k]
                return srv.submit_device(d_fields[k].ptr, r.w, r.h, r.seeds, d_times[i].ptr)
            for i in range(4):
                srv.wait(submit_dev(i))
            dev = _serve_runs(srv, submit_dev, n_requests, 5051, loads)
        dev.pop("results")
        # the last request of every distinct field, after the last load
        last = {pick[i]: i for i in range(n_requests)}
        served = [d_times[last[k]].download() for k in range(n_distinct)]
        worst = max(rel_diff(served[k], plain[k]) for k in range(n_distinct))
        dev.update({"plain_solves_per_s": round(1.0 / t_plain_dev, 1),
                    "speedup_vs_plain": round(dev["saturated_solves_per_s"] * t_plain_dev, 2),
                    "max_rel_diff_vs_plain": worst})
        out["device_resident"] = dev
        for d in d_fields + d_plain + d_times:
            d.free()
        # ---- host buffers: PCIe in and out ----
        for r in insts[:4]:
            ctx.solve(r.cost, r.seeds)
        t0 = time.perf_counter()
        for r in insts:
            ctx.solve(r.cost, r.seeds)
        t_plain_host = (time.perf_counter() - t0) / n_distinct
    hpick = [i % n_distinct for i in range(n_host_requests)]
    with g.Server(0, **serve_cfg) as srv:
        costs = [r.cost for r in insts]
        for i in range(4):
            srv.wait(srv.submit(costs[i], insts[i].seeds))
        host = _serve_runs(srv, lambda i: srv.submit(costs[hpick[i]], insts[hpick[i]].seeds),
                           n_host_requests, 5052, (0.8,))
    worst = max(rel_diff(t, plain[hpick[i]]) for i, (t, _) in enumerate(host.pop("results")))
    host.update({"plain_solves_per_s": round(1.0 / t_plain_host, 1),
                 "speedup_vs_plain": round(host["saturated_solves_per_s"] * t_plain_host, 2),
                 "max_rel_diff_vs_plain": worst})
    out["host_buffers"] = host
    # ---- accuracy against the FP64 FMM, and the CPU leg ----
    refs64 = _oracle_map(lambda r: oracle.fmm_f64(r.cost.astype(np.float64), r.seeds), insts)
    out["accuracy"] = {"served": accuracy_vs_fp64(served, refs64),
                       "plain": accuracy_vs_fp64(plain, refs64),
                       "reference": f"oracle FP64 FMM of each of the {n_distinct} distinct fields; "
                                    "served = the last request of each field in HBM serving"}
    out["linf_rel_vs_fp64"] = out["accuracy"]["served"]["linf_rel_vs_fp64"]
    del refs64, served, plain
    cores = host_cores()
    t0 = time.perf_counter()
    _oracle_map(lambda r: oracle.fmm_f32(r.cost, r.seeds), insts)
    t_cpu = time.perf_counter() - t0
    out["cpu_baseline"] = _with_host({"value": round(n_distinct / t_cpu, 2), "unit": "solves/s",
                                      "cores": cores, "kind": "port",
                                      "sample": f"oracle FP32 FMM over the {n_distinct} distinct "
                                                f"fields, {cores} threads, {t_cpu:.1f} s"})
    return out


def contrast_refs_async(g, pool):
    """The 4096^2 contrast field (rings at speed 1e-4) and, on a worker thread pinned to the
    last allowed core, its FP64 FMM and FP32 FIM (the FIM alone takes about 110 s on one
    core): started early so it overlaps the GPU legs."""
    from ghcm_amd import workloads
    r = workloads.contrast_4096()

    def work():
        import oracle
        try:
            os.sched_setaffinity(0, {max(os.sched_getaffinity(0))})   # this thread only
        except OSError:
            pass
        t0 = time.perf_counter()
        c64 = oracle.fmm_f64(1.0 / r.speed, r.seeds)
        cfim, _ = oracle.fim_f32(g.cost_from_speed(r.speed), r.seeds, 1e-5)
        return c64, cfim, time.perf_counter() - t0

    return r, pool.submit(work)


def bench_precision(g, cfg: dict, speed, seeds, ref64, ref_fim, contrast=None) -> dict:
    """The 4096^2 smooth field stored as FP32 cost and FP32/FP16/BF16 speed (times and updates
    in FP32), and the 4096^2 contrast field (rings at speed 1e-4) as FP32/FP16/BF16 speed: L_inf
    and L1 of each against the FP64 FMM and against the oracle's FP32 FIM (minimum change
    1e-5), beside the FIM's own error against the FP64 FMM."""
    import numpy as np

    import oracle

    def metrics(t, r64, rfim):
        m = oracle.error_metrics(t, r64)
        d = {"linf": m["linf"], "l1": m["l1"], "linf_rel": m["linf"] / m["max_ref"],
             "finiteness_mismatch": m["finiteness_mismatch"]}
        if rfim is not None:
            mf = oracle.error_metrics(t, rfim.astype(np.float64))
            d.update({"linf_vs_fim32": mf["linf"], "l1_vs_fim32": mf["l1"],
                      "linf_rel_vs_fim32": mf["linf"] / m["max_ref"]})
        return d

    def stored(sp):
        return [("speed_f32", sp.astype(np.float32), g.FIELD_SPEED_F32),
                ("speed_f16", sp.astype(np.float16), g.FIELD_SPEED_F16),
                ("speed_bf16", g.to_bf16(sp.astype(np.float32)), g.FIELD_SPEED_BF16)]

    out = {"reference": "oracle FP64 FMM; oracle FP32 FIM (minimum change 1e-5)",
           "tolerance_rel": 1e-4}
    with g.Context(0, **cfg) as ctx:
        sm = {}
        for name, field, kind in [("cost_f32", g.cost_from_speed(speed), g.FIELD_COST_F32)] + stored(speed):
            t, _ = ctx.solve(field, seeds, kind=kind)
            sm[name] = metrics(t, ref64, ref_fim)
        sm["fim32_vs_fp64"] = metrics(ref_fim, ref64, None)
        out["smooth_4096"] = sm
        if contrast is not None:
            r, fut = contrast
            c64, cfim, t_ref = fut.result()
            cf = {"workload": "4096^2, seven ring walls at speed 1e-4 on speed 1 (the 256^2 geometry "
                              f"x16), source at the centre; FP64 FMM and FP32 FIM {t_ref:.0f} s on one core"}
            for name, field, kind in stored(r.speed):
                t, _ = ctx.solve(field, r.seeds, kind=kind)
                cf[name] = metrics(t, c64, cfim)
            cf["fim32_vs_fp64"] = metrics(cfim, c64, None)
            out["contrast_4096"] = cf
    return out


def bench_reference_kernels(g, ctx, d_cost, d_time, seeds, steps: int) -> dict:
    """The 4096^2 solve with the reference's own in-tile algorithms (the systolic kernel at
    its default 8x8 blocks and at 32x8, the FIM kernel at 16x16), each with the reference's
    default wave size (n = CUs) and in-tile relaxation until convergence: time per solve and
    the tile kernel's roofline (algorithmic bytes per launch / average launch, HIP events)."""
    import numpy as np

    saved = {k: getattr(ctx.config, k) for k in ("tile_w", "tile_h", "kernel", "n_active", "inner_iter_cap")}
    out = {}
    ref = None
    for name, kern, (tw, th) in (("systolic_8x8", g.KERNEL_SYSTOLIC, (8, 8)),
                                 ("systolic_32x8", g.KERNEL_SYSTOLIC, (32, 8)),
                                 ("fim_16x16", g.KERNEL_FIM, (16, 16))):
        ctx.set_config(tile_w=tw, tile_h=th, kernel=kern, n_active=0, inner_iter_cap=0)

        def step():
            return ctx.solve_device(d_cost.ptr, g.FIELD_COST_F32, W, H, seeds, d_time.ptr)

        step()
        t0 = time.perf_counter()
        sts = [step() for _ in range(steps)]
        ms = (time.perf_counter() - t0) * 1e3 / steps
        t = d_time.download()
        ctx.set_config(profile_kernels=1)
        psts = [step() for _ in range(steps)]
        ctx.set_config(profile_kernels=0)
        launches = sum(x["solve_launches"] for x in psts)
        kern_us = sum(x["solve_kernel_us"] for x in psts)
        visits = sum(x["tile_visits"] for x in psts)
        bpl = visits * sts[0]["bytes_per_visit"] / max(launches, 1)
        avg_us = kern_us / max(launches, 1)
        achieved = bpl / (avg_us * 1e-6) / 1e9
        if ref is None:
            ref = t
        out[name] = {"ms_per_solve": round(ms, 3), "mcells_per_s": round(W * H / (ms / 1e3) / 1e6, 1),
                     "n_active": sts[0]["n_active"], "waves": sts[0]["waves"],
                     "tile_visits": sts[0]["tile_visits"],
                     "passes_per_visit": round(sts[0]["inner_iterations"] / max(1, sts[0]["tile_visits"]), 2),
                     "max_abs_diff_vs_first": float(np.max(np.abs(t - ref))),
                     "roofline": {"bound": "hbm", "kernel": "k_solve", "achieved": round(achieved, 2),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                                  "bytes_per_visit": sts[0]["bytes_per_visit"], "avg_launch_us": round(avg_us, 3),
                                  "launches_per_step": launches / steps}}
        log(f"bench: reference kernel {name}: {ms:.2f} ms per solve, frac {achieved / HBM_PEAK_GBS:.4f}")
    ctx.set_config(**saved)
    return out


def bench_launch_modes(g, ctx, d_cost, d_time, seeds, ref_t, steps: int) -> dict:
    """The headline 4096^2 solve, same configuration, in the boundary-free wave loops
    (no wave-boundary kernel): launch_mode 4 solves every queued tile each wave, 5 makes the
    top-n cut inside the tile launch.  Time per solve, waves, visits, the tile launch's
    average duration (HIP events) and the largest difference from the headline result."""
    import numpy as np

    saved = ctx.config.launch_mode
    out = {}
    for mode in (4, 5):
        ctx.set_config(launch_mode=mode)

        def step():
            return ctx.solve_device(d_cost.ptr, g.FIELD_COST_F32, W, H, seeds, d_time.ptr)

        for _ in range(2):
            step()
        t0 = time.perf_counter()
        sts = [step() for _ in range(steps)]
        ms = (time.perf_counter() - t0) * 1e3 / steps
        t = d_time.download()
        ctx.set_config(profile_kernels=1)
        psts = [step() for _ in range(3)]
        ctx.set_config(profile_kernels=0)
        launches = sum(x["solve_launches"] for x in psts)
        kern_us = sum(x["solve_kernel_us"] for x in psts)
        fin = np.isfinite(ref_t)
        out[f"launch_mode_{mode}"] = {
            "ms_per_solve": round(ms, 4), "mcells_per_s": round(W * H / (ms / 1e3) / 1e6, 1),
            "waves": sts[0]["waves"], "tile_visits": sts[0]["tile_visits"],
            "avg_launch_us": round(kern_us / max(launches, 1), 3), "launches_per_step": launches / 3,
            "max_abs_diff_vs_headline": float(np.max(np.abs(t[fin] - ref_t[fin]))),
            "finiteness_mismatch": int(np.count_nonzero(np.isfinite(t) != fin))}
        log(f"bench: launch_mode {mode}: {ms:.3f} ms per solve")
    ctx.set_config(launch_mode=saved)
    return out


def bench_volume(g, steps: int, cpu: bool) -> dict:
    """The 3-D (K = 3) solve on the 256^3 smooth volume: Mcells/s, the 3-D tile kernel's
    roofline, parity against the oracle's FP32 and FP64 3-D FMM, the oracle timed on one core."""
    import numpy as np

    from ghcm_amd import workloads
    t0 = time.perf_counter()
    speed, seeds = workloads.volume_256()
    cost = (np.float32(1.0) / speed).astype(np.float32)
    d, h, w = cost.shape
    log(f"bench: volume {w}x{h}x{d} generated in {time.perf_counter() - t0:.1f} s")
    out = {"workload": f"3-D {w}x{h}x{d} FP32 cost, smooth speed [0.1,10] sigma 8, "
                       f"{len(seeds)} point sources (seed 7)"}

    def run(kernel_cfg: dict, n_active: int):
        with g.Context(0, n_active=n_active, **kernel_cfg) as ctx:
            dc = ctx.to_device(cost)
            dt = ctx.alloc(cost.shape, np.float32)

            def step():
                return ctx.solve3d_device(dc.ptr, g.FIELD_COST_F32, w, h, d, seeds, dt.ptr)

            step()
            sts = []
            t0 = time.perf_counter()
            for _ in range(steps):
                sts.append(step())
            ms = (time.perf_counter() - t0) * 1e3 / steps
            t_gpu = dt.download()
            ctx.set_config(profile_kernels=1)
            psts = [step() for _ in range(steps)]
            ctx.set_config(profile_kernels=0)
        st = sts[0]
        launches = sum(x["solve_launches"] for x in psts)
        kern_us = sum(x["solve_kernel_us"] for x in psts)
        visits = sum(x["tile_visits"] for x in psts)
        bpl = visits * st["bytes_per_visit"] / max(launches, 1)
        avg_us = kern_us / max(launches, 1)
        achieved = bpl / (avg_us * 1e-6) / 1e9
        tile_cells = st["tile_w"] * st["tile_h"] * st["tile_d"]
        return t_gpu, {"tile": f"{st['tile_w']}x{st['tile_h']}x{st['tile_d']}",
                       "mcells_per_s": round(w * h * d / (ms / 1e3) / 1e6, 2), "ms_per_solve": round(ms, 4),
                       "n_active": n_active, "waves": st["waves"], "tile_visits": st["tile_visits"],
                       "rounds_per_visit": round(st["inner_iterations"] / max(st["tile_visits"], 1), 3),
                       "work_amplification": round(st["tile_visits"] * tile_cells / (w * h * d), 3),
                       "roofline": {"bound": "hbm", "achieved": round(achieved, 2),
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": round(achieved / HBM_PEAK_GBS, 5),
                                    "bytes_per_visit": st["bytes_per_visit"],
                                    "avg_launch_us": round(avg_us, 3),
                                    "launches_per_step": launches / steps}}

    # The octant-sweep kernel (64x8x8 tiles, one round per visit) is the measured path; the
    # in-tile-pass kernel (8x8x8 tiles) is timed beside it.
    t_gpu, res = run(dict(kernel=g.KERNEL_SWEEP, tile_w=64, tile_h=64, inner_iter_cap=1), 2048)
    res["kernel"] = "sweep (eight octant sweeps per round)"
    res["roofline"]["kernel"] = "k_solve3d_sweep"
    out.update(res)
    _, res_p = run(dict(kernel=g.KERNEL_SYSTOLIC), 4096)
    res_p["roofline"]["kernel"] = "k_solve3d"
    out["passes_8x8x8"] = res_p
    if cpu:
        import oracle
        t0 = time.perf_counter()
        ref32 = oracle.fmm3d_f32(cost, seeds)
        t_cpu = time.perf_counter() - t0
        ref64 = oracle.fmm3d_f64(cost.astype(np.float64), seeds)
        fin = np.isfinite(ref64)
        e32 = float(np.max(np.abs(t_gpu[fin].astype(np.float64) - ref32[fin])))
        m = oracle.error_metrics(t_gpu, ref64)
        out["cpu_baseline"] = {"value": round(w * h * d / t_cpu / 1e6, 3), "unit": "Mcells/s",
                               "cores": 1, "kind": "port",
                               "sample": f"full {w}x{h}x{d} volume, oracle FP32 3-D FMM, 1 trial, "
                                         f"{t_cpu:.1f} s"}
        # The FP32 oracle restates the reference's generic K-axis update, whose unshifted
        # quadratic cancels in FP32; its own error against FP64 is reported beside ours.
        e_ref = float(np.max(np.abs(ref32[fin].astype(np.float64) - ref64[fin])))
        out["parity"] = {"linf_vs_fp64": m["linf"], "l1_vs_fp64": m["l1"],
                         "linf_rel_fp64": m["linf"] / m["max_ref"],
                         "finiteness_mismatch": m["finiteness_mismatch"], "tolerance_rel": 1e-4,
                         "linf_rel_vs_cpu_fmm3d_f32": e32 / m["max_ref"],
                         "cpu_fmm3d_f32_linf_rel_vs_fp64": e_ref / m["max_ref"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tile", default="64x64")
    ap.add_argument("--kernel", default="sweep", choices=["systolic", "fim", "sweep"])
    ap.add_argument("--n-active", type=int, default=512)
    ap.add_argument("--inner-cap", type=int, default=1, help="in-tile rounds per visit (0 = converge)")
    ap.add_argument("--waves-per-poll", type=int, default=4)
    ap.add_argument("--launch-mode", type=int, default=0,
                    help="0 per-wave launches, 1 persistent, 2 graph replay, 3 fused boundary, 4 boundary-free, "
                         "5 boundary-free with the top-n cut made by the tile launch")
    ap.add_argument("--cpu-trials", type=int, default=5)
    ap.add_argument("--no-precision", action="store_true", help="skip the storage-precision report")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solve line")
    ap.add_argument("--no-batch-autotune", action="store_true",
                    help="solve the batch with the configuration the autotuner picks on MI355X "
                         "(64x64 sweep, n = 65,536) without running the search")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving line")
    ap.add_argument("--no-volume", action="store_true", help="skip the 3-D volume line")
    ap.add_argument("--no-launch-modes", action="store_true",
                    help="skip the headline solve in the boundary-free wave loops (launch_mode 4 and 5)")
    ap.add_argument("--no-reference-kernels", action="store_true",
                    help="skip the 4096^2 solve with the systolic and FIM tile kernels")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this size first (rocprofv3 --pmc scaling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the gHCM path is a single-"
              "device solve; sharding across GPUs is designed but not built). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import numpy as np

    import ghcm_amd as g

    tw, th = (int(v) for v in args.tile.split("x"))
    kernel = {"fim": g.KERNEL_FIM, "systolic": g.KERNEL_SYSTOLIC, "sweep": g.KERNEL_SWEEP}[args.kernel]
    log(f"bench: generating {W}x{H} workload")
    speed = g.gen_smooth_speed(W, H, FIELD_SEED, 32.0, 0.1, 10.0)
    seeds = g.gen_points(W, H, POINT_SEED, 4)
    cost = g.cost_from_speed(speed)

    ctx = g.Context(0, tile_w=tw, tile_h=th, kernel=kernel, n_active=args.n_active,
                    inner_iter_cap=args.inner_cap, waves_per_poll=args.waves_per_poll,
                    launch_mode=args.launch_mode, profile_kernels=0)
    info = g.device_info(0)
    d_cost = ctx.to_device(cost)
    d_time = ctx.alloc(cost.shape, np.float32)
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib << 20
        src, dst = ctx.alloc((n,), np.uint8), ctx.alloc((n,), np.uint8)
        g.lib().ghcm_dev_copy(ctx._h, dst.ptr, src.ptr, n)
        src.free()
        dst.free()

    def step():
        return ctx.solve_device(d_cost.ptr, g.FIELD_COST_F32, W, H, seeds, d_time.ptr)

    for i in range(args.warmup):
        st = step()
        log(f"bench: warmup {i + 1}/{args.warmup} {st['duration_device_us'] / 1e3:.2f} ms")

    # Timed region: each solve_device call returns after the solver stream drained.
    sts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sts.append(step())
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = W * H / (ms_per_step / 1e3) / 1e6
    log(f"bench: {args.steps} steps, {ms_per_step:.3f} ms/step, {value:.1f} Mcells/s")

    t_gpu = d_time.download()
    # The same solve through the host-buffer entry point (field in, times out over PCIe):
    # reported beside `value`, never as it.
    hts = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.solve(cost, seeds)
        hts.append(time.perf_counter() - t0)
    host_ms = statistics.median(hts) * 1e3
    # Roofline of the tile-solve kernel over the same number of instrumented steps.
    ctx.set_config(profile_kernels=1)
    psts = [step() for _ in range(args.steps)]
    ctx.set_config(profile_kernels=0)
    launches = sum(s["solve_launches"] for s in psts)
    kern_us = sum(s["solve_kernel_us"] for s in psts)
    visits = sum(s["tile_visits"] for s in psts)
    bpv = sts[0]["bytes_per_visit"]
    bytes_per_launch = visits * bpv / max(launches, 1)
    avg_launch_us = kern_us / max(launches, 1)
    achieved = bytes_per_launch / (avg_launch_us * 1e-6) / 1e9
    traffic = read_traffic()
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None if traffic is None else round(traffic, 1),
            "kernel": "k_solve_sweep" if kernel == g.KERNEL_SWEEP else "k_solve",
            "avg_launch_us": round(avg_launch_us, 3),
            "bytes_per_launch": round(bytes_per_launch, 1),
            "launches_per_step": launches / args.steps,
            "ms_per_step_instrumented": round(
                statistics.mean(s["duration_total_us"] for s in psts) / 1e3, 4)}

    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mcells/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded smooth log-normal speed field)",
           "config": {"workload": f"plain gHCM {W}x{H} FP32 cost, smooth speed [0.1,10] "
                                  f"sigma 32, 4 point sources",
                      "tile": f"{tw}x{th}", "kernel": args.kernel,
                      "n_active": args.n_active, "inner_rounds_cap": args.inner_cap,
                      "launch_mode": args.launch_mode, "scheduler": "device_topn",
                      "fim_minimum_change": 1e-5, "parallelism": "single_gpu",
                      "device": info["arch"]},
           "solve": {"waves": sts[0]["waves"], "tile_visits": sts[0]["tile_visits"],
                     "work_amplification": round(sts[0]["tile_visits"] * tw * th / (W * H), 3),
                     "inner_passes_per_visit": round(sts[0]["inner_iterations"] /
                                                     max(1, sts[0]["tile_visits"]), 2),
                     "device_ms": round(statistics.median(s["duration_device_us"] for s in sts) / 1e3, 4)},
           "roofline": roof,
           "host_buffers": {"ms_per_solve": round(host_ms, 3),
                            "mcells_per_s": round(W * H / (host_ms / 1e3) / 1e6, 1),
                            "note": "ghcm_solve with the field and times in host memory "
                                    "(128 MB over PCIe per solve)"}}

    if not args.no_reference_kernels:
        out["reference_kernels"] = bench_reference_kernels(g, ctx, d_cost, d_time, seeds, 3)
    if not args.no_launch_modes and kernel == g.KERNEL_SWEEP and args.launch_mode == 0:
        out["launch_modes"] = bench_launch_modes(g, ctx, d_cost, d_time, seeds, t_gpu, args.steps)

    plain_cfg = dict(tile_w=tw, tile_h=th, kernel=kernel, n_active=args.n_active,
                     inner_iter_cap=args.inner_cap, waves_per_poll=args.waves_per_poll,
                     launch_mode=args.launch_mode)
    # The contrast field's oracle solutions run on one host core while the GPU legs below run.
    import concurrent.futures as cf
    bg = cf.ThreadPoolExecutor(max_workers=1)
    contrast = None if args.no_precision else contrast_refs_async(g, bg)
    if not args.no_batch:
        log("bench: batched solves")
        out["batched"] = bench_batch(g, plain_cfg, 3, autotune=not args.no_batch_autotune)
        log(f"bench: batched {out['batched']['solves_per_s']} solves/s, "
            f"{out['batched']['speedup_vs_plain']}x plain")
    if not args.no_serve:
        log("bench: serving")
        # A serving loop runs many requests per wave: without the batch autotuner's choice it
        # selects every queued tile (n = 65,536), as the autotuner does for the batch.
        scfg = dict(plain_cfg, n_active=65536)
        if "autotuned" in out.get("batched", {}):
            scfg.update(out["batched"]["autotuned"])
        out["serving"] = bench_serving(g, scfg)
        sv = out["serving"]
        l80 = sv["device_resident"]["loads"]["0.80"]
        log(f"bench: serving (HBM-resident) {sv['device_resident']['saturated_solves_per_s']} "
            f"solves/s, {sv['device_resident']['speedup_vs_plain']}x plain, p50 "
            f"{l80['p50_ms']} ms p99 {l80['p99_ms']} ms at 80% load; "
            f"host buffers {sv['host_buffers']['saturated_solves_per_s']} solves/s, "
            f"{sv['host_buffers']['speedup_vs_plain']}x plain")

    if not args.no_volume:
        log("bench: 3-D volume")
        out["volume"] = bench_volume(g, 3, not args.no_cpu_baseline)
        log(f"bench: volume {out['volume']['mcells_per_s']} Mcells/s, "
            f"{out['volume']['ms_per_solve']} ms per solve")

    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline ({args.cpu_trials} trials of the oracle FMM, 1 of the FIM)")
        ref, cb = cpu_baseline(cost, seeds, args.cpu_trials)
        out["cpu_baseline"] = cb
        import oracle
        t0 = time.perf_counter()
        ref_fim, _ = oracle.fim_f32(cost, seeds, 1e-5)
        t_fim = time.perf_counter() - t0
        out["cpu_baseline_fim"] = _with_host({"value": round(W * H / t_fim / 1e6, 3), "unit": "Mcells/s",
                                              "cores": 1, "kind": "port",
                                              "sample": f"full {W}x{H} workload, oracle FP32 FIM "
                                                        f"(minimum change 1e-5), 1 trial, {t_fim:.1f} s"})
        err = float(np.max(np.abs(t_gpu.astype(np.float64) - ref)))
        fin = np.isfinite(ref_fim)
        err_fim = float(np.max(np.abs(t_gpu[fin].astype(np.float64) - ref_fim[fin])))
        out["parity"] = {"linf_vs_cpu_fmm": err, "linf_rel": err / float(ref.max()),
                         "linf_vs_cpu_fim": err_fim, "linf_rel_fim": err_fim / float(ref.max()),
                         "tolerance_rel": 1e-4}
    else:
        out["cpu_baseline"] = None
        ref_fim = None
    if not args.no_precision:
        import oracle
        log("bench: storage precision vs the FP64 FMM and the FP32 FIM")
        ref64 = oracle.fmm_f64(1.0 / speed, seeds)
        if ref_fim is None:
            ref_fim, _ = oracle.fim_f32(cost, seeds, 1e-5)
        out["precision"] = bench_precision(g, plain_cfg, speed, seeds, ref64, ref_fim, contrast)
    ctx.close()
    bg.shutdown(wait=True)
    out["host"] = host_description()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
