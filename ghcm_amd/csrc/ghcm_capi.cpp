// ghcm_capi.cpp -- C-ABI of libghcm: devices, contexts, and the two wave schedulers.
//
// Replaces src/util_opencl.{h,cpp} (device enumeration/info, error strings, cl_do) and the
// host half of eikonal_hcm_opencl_ex (src/eikonal_hcm.cpp:186-529).  Nothing here exits
// or throws across the C boundary: every entry point returns a ghcm_status and records a
// message retrievable with ghcm_last_error().
#include "ghcm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <new>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "ghcm_dev.h"
#include "ghcm_io.h"
#include "kernels.h"
#include "tile_geom.h"

using namespace ghcm;

#include "capi_common.h"

using ghcm_capi::DevBuf;
using ghcm_capi::PinnedBuf;
using ghcm_capi::bytes_per_visit;
using ghcm_capi::bytes_per_visit3d;
using ghcm_capi::fail;
using ghcm_capi::kind_bytes;
using ghcm_capi::no_throw;
using ghcm_capi::valid_kind;

namespace {
thread_local std::string g_last_error;
// Seeds the allocation-free solve accepts (its pinned seed list is sized when a context is
// created).
constexpr uint32_t kNoAllocSeeds = 4096;
}  // namespace

int ghcm_capi::fail(int status, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

struct ghcm_ctx {
    int ordinal = 0;
    ghcm_config cfg{};
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    hipEvent_t ev_sync = nullptr;       // polled in place of stream synchronisation
    std::vector<hipEvent_t> ev_prof;    // profile_kernels: begin/end pairs per queued wave
    std::vector<hipEvent_t> ev_slot;    // device scheduler: one per chunk in flight
    // launch_mode 2: one chunk of waves captured as a HIP graph, reused while the solve
    // arguments (buffers, sizes, n) and grid stay the same.
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;
    // device scratch
    DevBuf inst, tinst, qwave, key, mask, tmin, list, ctrl, seeds, jobs, status, tstat, totals, sel;
    DevBuf edge, cedge;         // sweep kernel: edge columns of every tile (time, cost)
    DevBuf field_io, time_io;   // staging for the host-buffer entry point
    PinnedBuf io_stage[2];      // pinned parts of host-buffer copies (ping-pong)
    hipEvent_t ev_io[2] = {nullptr, nullptr};
    PinnedBuf h_ctrl, h_seeds, h_inst, h_jobs, h_status;
    PinnedBuf rec;                       // HostRec written by the wave loop
    uint32_t solve_seq = 0;              // tags the totals of each solve
};

// ---------------------------------------------------------------------------------------
// parallel host copies
// ---------------------------------------------------------------------------------------

namespace {

// Worker pool of par_memcpy.  Never destroyed: its threads wait on a condition variable for
// the life of the process.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* pool = new CopyPool();
        return *pool;
    }
    size_t workers() const { return threads_.size(); }
    // Runs parts [1, parts) on the workers and part 0 on the caller; returns when all are done.
    void run(size_t parts, const std::function<void(size_t)>& part) {
        struct Batch {
            std::atomic<size_t> left;
        } batch{{parts - 1}};
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = 1; i < parts; ++i)
                queue_.push_back([&part, &batch, i] {
                    part(i);
                    batch.left.fetch_sub(1, std::memory_order_release);
                });
        }
        cv_.notify_all();
        part(0);
        // Help with queued parts (ours or another caller's) instead of only waiting.
        while (batch.left.load(std::memory_order_acquire) != 0) {
            std::function<void()> job;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (!queue_.empty()) {
                    job = std::move(queue_.front());
                    queue_.pop_front();
                }
            }
            if (job) job();
            else std::this_thread::yield();
        }
    }

private:
    CopyPool() {
        const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
        const size_t n = std::min<size_t>(7, hw > 1 ? hw - 1 : 0);
        for (size_t i = 0; i < n; ++i) {
            try {
                threads_.emplace_back([this] { loop(); });
                threads_.back().detach();
            } catch (...) {
                break;
            }
        }
    }
    void loop() {
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return !queue_.empty(); });
                job = std::move(queue_.front());
                queue_.pop_front();
            }
            job();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> queue_;
    std::vector<std::thread> threads_;
};

}  // namespace

void ghcm_capi::par_memcpy(void* dst, const void* src, size_t n) {
    constexpr size_t kPart = size_t(2) << 20;
    CopyPool& pool = CopyPool::get();
    const size_t parts = std::min<size_t>(n / kPart, pool.workers() + 1);
    if (parts < 2) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + parts - 1) / parts + kPart - 1) / kPart * kPart;
    auto* d = static_cast<char*>(dst);
    const auto* sp = static_cast<const char*>(src);
    pool.run(parts, [d, sp, n, per](size_t i) {
        const size_t off = i * per;
        if (off < n) std::memcpy(d + off, sp + off, std::min(per, n - off));
    });
}

// ---------------------------------------------------------------------------------------
// library / device
// ---------------------------------------------------------------------------------------

extern "C" int ghcm_abi_version(void) { return GHCM_ABI_VERSION; }

extern "C" const char* ghcm_status_string(int s) {
    switch (s) {
        case GHCM_OK: return "ok";
        case GHCM_ERR_INVALID_ARG: return "invalid argument";
        case GHCM_ERR_NO_DEVICE: return "no such device";
        case GHCM_ERR_HIP: return "HIP runtime error";
        case GHCM_ERR_OUT_OF_MEMORY: return "out of memory";
        case GHCM_ERR_NOT_CONVERGED: return "did not converge within the wave cap";
        case GHCM_ERR_UNSUPPORTED: return "no kernel built for this configuration";
        case GHCM_ERR_OVER_BUDGET: return "request exceeds the HBM budget";
        case GHCM_ERR_NOT_FOUND: return "unknown ticket";
        case GHCM_ERR_STOPPED: return "server stopped";
        case GHCM_ERR_TIMEOUT: return "wait timed out";
    }
    return "unknown status";
}

extern "C" const char* ghcm_last_error(void) { return g_last_error.c_str(); }

// HcmConfig defaults (src/eikonal.h:12-27): block 8, overloads 1, epsilon 1e-5,
// cache Both, Systolic.
extern "C" void ghcm_config_default(ghcm_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->tile_w = 8;
    c->tile_h = 8;
    c->workers_overload = 1.0f;
    c->compute_unit_overload = 1.0f;
    c->n_active = 0;
    c->min_change = 1e-5f;
    c->cache_mode = GHCM_CACHE_BOTH;
    c->kernel = GHCM_KERNEL_SYSTOLIC;
    c->scheduler = GHCM_SCHED_DEVICE_TOPN;
    c->inner_iter_cap = 0;
    c->max_waves = 0;
    c->waves_per_poll = 0;
    c->profile_kernels = 0;
    c->launch_mode = 0;
}

extern "C" int ghcm_device_count(int* count) {
    if (!count) return fail(GHCM_ERR_INVALID_ARG, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(GHCM_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = n;
    return GHCM_OK;
}

extern "C" int ghcm_device_info_get(int ordinal, ghcm_device_info* out) {
    if (!out) return fail(GHCM_ERR_INVALID_ARG, "out is null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ordinal < 0 || ordinal >= n)
        return fail(GHCM_ERR_NO_DEVICE, "device %d not present (%d devices)", ordinal, n);
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, ordinal));
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", p.name);
    std::snprintf(out->arch, sizeof out->arch, "%.*s", (int)sizeof out->arch - 1, p.gcnArchName);
    out->ordinal = ordinal;
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->max_workgroup = p.maxThreadsPerBlock;
    out->lds_bytes = p.sharedMemPerBlock;
    out->global_mem_bytes = p.totalGlobalMem;
    out->pci_bus_id = p.pciBusID;
    return GHCM_OK;
}

// ---------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------

int ghcm_capi::check_config(const ghcm_config* c) {
    if (!c) return fail(GHCM_ERR_INVALID_ARG, "config is null");
    if (!(c->min_change >= 0.0f) || !std::isfinite(c->min_change))
        return fail(GHCM_ERR_INVALID_ARG, "min_change must be finite and >= 0");
    if (c->kernel != GHCM_KERNEL_SYSTOLIC && c->kernel != GHCM_KERNEL_FIM &&
        c->kernel != GHCM_KERNEL_SWEEP)
        return fail(GHCM_ERR_INVALID_ARG, "unknown kernel %d", c->kernel);
    if (c->scheduler != GHCM_SCHED_DEVICE_TOPN && c->scheduler != GHCM_SCHED_HOST_PQ)
        return fail(GHCM_ERR_INVALID_ARG, "unknown scheduler %d", c->scheduler);
    if (c->cache_mode < GHCM_CACHE_NONE || c->cache_mode > GHCM_CACHE_BOTH)
        return fail(GHCM_ERR_INVALID_ARG, "unknown cache mode %d", c->cache_mode);
    if (!(c->compute_unit_overload > 0.0f) || !(c->workers_overload > 0.0f))
        return fail(GHCM_ERR_INVALID_ARG, "overloads must be > 0");
    for (int k = 0; k < 4; ++k)
        if (!get_solve_kernel((int)c->tile_w, (int)c->tile_h, k, c->kernel, false).launch)
            return fail(GHCM_ERR_UNSUPPORTED, "no kernel %d for tile %ux%u", c->kernel, c->tile_w,
                        c->tile_h);
    return GHCM_OK;
}

extern "C" int ghcm_ctx_create(int ordinal, const ghcm_config* cfg, ghcm_ctx** out) {
    if (!out) return fail(GHCM_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    ghcm_config c;
    if (cfg) c = *cfg; else ghcm_config_default(&c);
    int rc = ghcm_capi::check_config(&c);
    if (rc) return rc;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ordinal < 0 || ordinal >= n)
        return fail(GHCM_ERR_NO_DEVICE, "device %d not present (%d devices)", ordinal, n);
    HIP_TRY(hipSetDevice(ordinal));
    ghcm_ctx* ctx = new (std::nothrow) ghcm_ctx();
    if (!ctx) return fail(GHCM_ERR_OUT_OF_MEMORY, "context allocation failed");
    ctx->ordinal = ordinal;
    ctx->cfg = c;
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, ordinal);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_begin);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_end);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_sync, hipEventDisableTiming);
    if (e != hipSuccess) {
        ghcm_ctx_destroy(ctx);
        return fail(GHCM_ERR_HIP, "context setup failed: %s", hipGetErrorString(e));
    }
    ctx->cus = p.multiProcessorCount;
    // Set-up for the allocation-free solve, done here so that entry point allocates nothing:
    // this device's address of the static scratch array (looked up once per device, with the
    // device current), and the pinned records every solve uses -- the progress record, one
    // instance entry and room for kNoAllocSeeds seeds.  They only grow later.
    e = noalloc_scratch_init(ordinal);
    if (e != hipSuccess) {
        ghcm_ctx_destroy(ctx);
        return fail(GHCM_ERR_HIP, "scratch address lookup failed: %s", hipGetErrorString(e));
    }
    int rc2 = ctx->rec.reserve(sizeof(HostRec));
    if (!rc2) rc2 = ctx->h_inst.reserve(sizeof(GhcmInst));
    if (!rc2) rc2 = ctx->h_seeds.reserve(3ull * 4 * kNoAllocSeeds);
    if (rc2) {
        ghcm_ctx_destroy(ctx);
        return rc2;
    }
    *out = ctx;
    return GHCM_OK;
}

extern "C" int ghcm_ctx_destroy(ghcm_ctx* ctx) {
    if (!ctx) return GHCM_OK;
    (void)hipSetDevice(ctx->ordinal);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (DevBuf* b : {&ctx->inst, &ctx->tinst, &ctx->qwave, &ctx->key, &ctx->mask, &ctx->tmin,
                      &ctx->list, &ctx->ctrl, &ctx->seeds, &ctx->jobs, &ctx->status,
                      &ctx->tstat, &ctx->totals, &ctx->sel, &ctx->edge, &ctx->cedge,
                      &ctx->field_io, &ctx->time_io})
        b->release();
    for (PinnedBuf* b : {&ctx->h_ctrl, &ctx->h_seeds, &ctx->h_inst, &ctx->h_jobs,
                         &ctx->h_status, &ctx->rec, &ctx->io_stage[0], &ctx->io_stage[1]})
        b->release();
    for (hipEvent_t ev : ctx->ev_io)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->ev_begin) (void)hipEventDestroy(ctx->ev_begin);
    if (ctx->ev_end) (void)hipEventDestroy(ctx->ev_end);
    if (ctx->ev_sync) (void)hipEventDestroy(ctx->ev_sync);
    for (hipEvent_t ev : ctx->ev_prof)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : ctx->ev_slot)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GHCM_OK;
}

extern "C" int ghcm_ctx_get_config(const ghcm_ctx* ctx, ghcm_config* out) {
    if (!ctx || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    *out = ctx->cfg;
    return GHCM_OK;
}

extern "C" int ghcm_ctx_set_config(ghcm_ctx* ctx, const ghcm_config* cfg) {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    int rc = ghcm_capi::check_config(cfg);
    if (rc) return rc;
    ctx->cfg = *cfg;
    return GHCM_OK;
}

// ---------------------------------------------------------------------------------------
// shared solve core
// ---------------------------------------------------------------------------------------

namespace {

struct HostInst {      // host view of one instance
    const void* d_field;
    float* d_time;
    uint32_t w, h;
    const uint32_t* seeds;  // (x, y) pairs, or (x, y, z) triples for a 3-D solve
    uint32_t n_seeds;
    uint32_t d = 1;         // depth of a 3-D instance
};


// One device wave loop over n_inst instances.  three_d: every instance is a W x H x D grid
// solved in 8x8x8 tiles by the 3-D tile kernel (device scheduler, per-wave launches or
// graph replay); the configuration's tile shape and in-tile style do not apply.
// noalloc: the allocation-free variant (ghcm_solve_device_noalloc): every device buffer is
// carved out of the device's static scratch array, the pinned host records must already
// hold what the solve needs (ghcm_ctx_create reserves them), and configurations that would
// allocate (graph capture, per-launch profiling events, the host scheduler's containers,
// the persistent and fused loops' extra records) are refused.
int solve_core_body(ghcm_ctx* ctx, int kind, const HostInst* hi, uint32_t n_inst, ghcm_stats* stats,
                    bool three_d, bool noalloc);

int solve_core(ghcm_ctx* ctx, int kind, const HostInst* hi, uint32_t n_inst, ghcm_stats* stats,
               bool three_d = false, bool noalloc = false) {
    return no_throw([&] { return solve_core_body(ctx, kind, hi, n_inst, stats, three_d, noalloc); });
}

// Byte offsets of the scheduler state of a T-tile solve inside the scratch array; each
// region starts on a 256-byte boundary; `end` is the bytes used.
struct ScratchLayout {
    size_t ctrl, inst, totals, qwave, key, mask, tmin, tstat, sel, list, edge, cedge, end;
};
// edge_h: rows of a tile's edge columns (the sweep kernel's TH), 0 without edge columns.
ScratchLayout scratch_layout(uint32_t T, uint32_t shard_cap, uint32_t edge_h = 0) {
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    ScratchLayout l{};
    size_t o = 0;
    l.ctrl = o;   o = up(o + sizeof(GhcmCtrl));
    l.inst = o;   o = up(o + sizeof(GhcmInst));
    l.totals = o; o = up(o + 3 * sizeof(unsigned long long));
    l.qwave = o;  o = up(o + 4ull * T);
    l.key = o;    o = up(o + 8ull * T);
    l.mask = o;   o = up(o + 8ull * T);
    l.tmin = o;   o = up(o + 4ull * T);
    l.tstat = o;  o = up(o + 12ull * T);
    l.sel = o;    o = up(o + 4ull * std::max<uint32_t>(1, T));
    l.list = o;   o = up(o + 12ull * kShards * shard_cap);
    l.edge = o;   o = up(o + 8ull * edge_h * T);
    l.cedge = o;  o = up(o + 8ull * edge_h * T);
    l.end = o;
    return l;
}

// Largest tile count whose scheduler state fits the scratch array.
uint32_t noalloc_max_tiles(uint32_t edge_h) {
    uint32_t lo = 1, hi = 1u << 24;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (scratch_layout(mid, (mid + kShards - 1) / kShards, edge_h).end <= kNoAllocScratchBytes) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

int solve_core_body(ghcm_ctx* ctx, int kind, const HostInst* hi, uint32_t n_inst, ghcm_stats* stats,
                    bool three_d, bool noalloc) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    const ghcm_config& cfg = ctx->cfg;
    if (!valid_kind(kind)) return fail(GHCM_ERR_INVALID_ARG, "unknown field kind %d", kind);
    if (n_inst == 0) return fail(GHCM_ERR_INVALID_ARG, "no instances");
    if (noalloc) {
        if (n_inst != 1 || three_d) return fail(GHCM_ERR_INVALID_ARG, "allocation-free solve: one 2-D instance");
        if (cfg.scheduler != GHCM_SCHED_DEVICE_TOPN || (cfg.launch_mode != 0 && cfg.launch_mode != 4) ||
            cfg.profile_kernels)
            return fail(GHCM_ERR_UNSUPPORTED, "allocation-free solve: device scheduler, launch_mode 0 or 4, "
                                              "no kernel profiling");
    }
    // Pinned host records: grown on demand, or (allocation-free solve) checked against what
    // the context reserved when it was created.
    auto pinned = [noalloc](PinnedBuf& b, size_t want, const char* what) -> int {
        if (!noalloc) return b.reserve(want);
        if (b.bytes < want)
            return fail(GHCM_ERR_INVALID_ARG, "allocation-free solve: %s needs %zu bytes, %zu reserved", what,
                        want, b.bytes);
        return GHCM_OK;
    };
    // 3-D solves: the sweep kernel's 64 x 8 x 8 tiles, or 8 x 8 x 8 tiles of in-tile passes.
    const bool sweep3 = three_d && cfg.kernel == GHCM_KERNEL_SWEEP;
    uint32_t tw = cfg.tile_w, th = cfg.tile_h, td = 1u;
    if (three_d) tile3d_shape(sweep3, tw, th, td);
    const bool host_sched = cfg.scheduler == GHCM_SCHED_HOST_PQ;
    if (three_d && host_sched)
        return fail(GHCM_ERR_UNSUPPORTED, "3-D solves run with the device scheduler only");
    uint32_t n_sel = cfg.n_active ? cfg.n_active
                                  : (uint32_t)std::max(1.0, std::floor((double)cfg.compute_unit_overload * ctx->cus));
    // Waves far wider than the chip (batches of many instances) take the higher-occupancy
    // form of the tile kernel where one is built.
    const bool wide = (uint64_t)n_sel >= 4ull * (uint64_t)std::max(1, ctx->cus);
    const SolveKernel sk = three_d ? get_solve3d_kernel(kind, sweep3)
                                   : get_solve_kernel((int)tw, (int)th, kind, cfg.kernel, host_sched, wide);
    const solve_launcher_t launch = sk.launch;
    if (!launch)
        return fail(GHCM_ERR_UNSUPPORTED, "no kernel for tile %ux%u kind %d kernel %d%s", tw, th,
                    kind, cfg.kernel, host_sched ? " (host scheduler)" : "");
    const uint32_t sd = three_d ? 3u : 2u;            // seed components

    // ---- instance table, seeds ----
    int rc = pinned(ctx->h_inst, sizeof(GhcmInst) * n_inst, "the instance table");
    if (rc) return rc;
    GhcmInst* hinst = ctx->h_inst.as<GhcmInst>();
    uint64_t tiles64 = 0, cells = 0, n_seeds_total = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const HostInst& x = hi[i];
        if (x.w == 0 || x.h == 0) return fail(GHCM_ERR_INVALID_ARG, "instance %u has zero size", i);
        if (x.w > GHCM_MAX_WIDTH)
            return fail(GHCM_ERR_INVALID_ARG, "instance %u width %u above %u", i, x.w, GHCM_MAX_WIDTH);
        if (!x.d_field || !x.d_time) return fail(GHCM_ERR_INVALID_ARG, "instance %u has a null buffer", i);
        if (x.n_seeds && !x.seeds) return fail(GHCM_ERR_INVALID_ARG, "instance %u seeds are null", i);
        const uint32_t dd = three_d ? x.d : 1u;
        if (dd == 0) return fail(GHCM_ERR_INVALID_ARG, "instance %u has zero depth", i);
        for (uint32_t s = 0; s < x.n_seeds; ++s) {
            const uint32_t sx = x.seeds[sd * s], sy = x.seeds[sd * s + 1], sz = three_d ? x.seeds[sd * s + 2] : 0u;
            if (sx >= x.w || sy >= x.h || sz >= dd)
                return fail(GHCM_ERR_INVALID_ARG, "instance %u seed %u (%u,%u,%u) outside %ux%ux%u", i, s,
                            sx, sy, sz, x.w, x.h, dd);
        }
        GhcmInst& g = hinst[i];
        g.field = x.d_field;
        g.time = x.d_time;
        g.w = x.w;
        g.h = x.h;
        g.tiles_x = tiles_along(x.w, tw);
        g.tiles_y = tiles_along(x.h, th);
        g.tile_base = (uint32_t)tiles64;
        g.d = dd;
        g.tiles_z = (dd + td - 1) / td;
        g.pad = 0;
        tiles64 += (uint64_t)g.tiles_x * g.tiles_y * g.tiles_z;
        if (tiles64 >= 0x7FFFFFFFull) return fail(GHCM_ERR_INVALID_ARG, "too many tiles");
        cells += (uint64_t)x.w * x.h * dd;
        n_seeds_total += x.n_seeds;
    }
    if (tiles64 >= 0x7FFFFFFFull) return fail(GHCM_ERR_INVALID_ARG, "too many tiles");
    const uint32_t T = (uint32_t)tiles64;

    HIP_TRY(hipSetDevice(ctx->ordinal));
    hipStream_t st = ctx->stream;
    const uint32_t shard_cap = (T + kShards - 1) / kShards;
    // The sweep kernels stage their halo columns from edge columns kept for every tile
    // (th * td rows each; a 3-D sweep tile's 64 rows).
    const uint32_t edge_h = (cfg.kernel == GHCM_KERNEL_SWEEP) ? th * td : 0u;
    float *d_edge = nullptr, *d_cedge = nullptr;
    // Device scheduler state: the context's grow-only buffers, or the scratch array.
    GhcmInst* d_inst = nullptr;
    uint32_t *d_tinst = nullptr, *d_qwave = nullptr, *d_key = nullptr, *d_mask = nullptr;
    uint32_t *d_list = nullptr, *d_tstat = nullptr, *d_sel = nullptr;
    float* d_tmin = nullptr;
    unsigned long long* d_totals = nullptr;
    GhcmCtrl* d_ctrl = nullptr;
    if (noalloc) {
        char* const base = static_cast<char*>(noalloc_scratch(ctx->ordinal));
        if (!base) return fail(GHCM_ERR_HIP, "allocation-free solve: no scratch address for device %d", ctx->ordinal);
        const ScratchLayout l = scratch_layout(T, shard_cap, edge_h);
        if (l.end > kNoAllocScratchBytes)
            return fail(GHCM_ERR_INVALID_ARG, "allocation-free solve: %u tiles of %ux%u need %zu bytes of "
                                              "scratch, above its %zu (at most %u tiles)", T, tw, th, l.end,
                        kNoAllocScratchBytes, noalloc_max_tiles(edge_h));
        if (edge_h) {
            d_edge = reinterpret_cast<float*>(base + l.edge);
            d_cedge = reinterpret_cast<float*>(base + l.cedge);
        }
        d_ctrl = reinterpret_cast<GhcmCtrl*>(base + l.ctrl);
        d_inst = reinterpret_cast<GhcmInst*>(base + l.inst);
        d_totals = reinterpret_cast<unsigned long long*>(base + l.totals);
        d_qwave = reinterpret_cast<uint32_t*>(base + l.qwave);
        d_key = reinterpret_cast<uint32_t*>(base + l.key);
        d_mask = reinterpret_cast<uint32_t*>(base + l.mask);
        d_tmin = reinterpret_cast<float*>(base + l.tmin);
        d_tstat = reinterpret_cast<uint32_t*>(base + l.tstat);
        d_sel = reinterpret_cast<uint32_t*>(base + l.sel);
        d_list = reinterpret_cast<uint32_t*>(base + l.list);
    } else {
        if ((rc = ctx->inst.reserve(sizeof(GhcmInst) * n_inst))) return rc;
        if ((rc = ctx->qwave.reserve(4ull * T))) return rc;
        // activation records: two parities, three when the boundary-free launch makes the
        // top-n cut (launch_mode 5 rotates them with the lists)
        if ((rc = ctx->key.reserve(12ull * T))) return rc;
        if ((rc = ctx->mask.reserve(12ull * T))) return rc;
        if ((rc = ctx->tmin.reserve(4ull * T))) return rc;
        // three lists: boundary-free waves rotate them (the other modes use two)
        if ((rc = ctx->list.reserve(12ull * kShards * shard_cap))) return rc;
        if ((rc = ctx->tstat.reserve(12ull * T))) return rc;
        if ((rc = ctx->totals.reserve(3 * sizeof(unsigned long long)))) return rc;
        if ((rc = ctx->ctrl.reserve(sizeof(GhcmCtrl)))) return rc;
        if (n_inst > 1 && (rc = ctx->tinst.reserve(4ull * T))) return rc;
        // jobs[]: the selected tiles plus the carried tiles a long selection hands to the tile
        // launch, so one word per tile
        if ((rc = ctx->sel.reserve(4ull * std::max<uint64_t>(1, T)))) return rc;
        if (ctx->sel.bytes < 4ull * T) return fail(GHCM_ERR_INVALID_ARG, "selection buffer below %u words", T);
        d_inst = ctx->inst.as<GhcmInst>();
        d_tinst = n_inst > 1 ? ctx->tinst.as<uint32_t>() : nullptr;
        d_qwave = ctx->qwave.as<uint32_t>();
        d_key = ctx->key.as<uint32_t>();
        d_mask = ctx->mask.as<uint32_t>();
        d_tmin = ctx->tmin.as<float>();
        d_list = ctx->list.as<uint32_t>();
        d_tstat = ctx->tstat.as<uint32_t>();
        d_totals = ctx->totals.as<unsigned long long>();
        d_ctrl = ctx->ctrl.as<GhcmCtrl>();
        d_sel = ctx->sel.as<uint32_t>();
        if (edge_h) {
            if ((rc = ctx->edge.reserve(8ull * edge_h * T))) return rc;
            if ((rc = ctx->cedge.reserve(8ull * edge_h * T))) return rc;
            d_edge = ctx->edge.as<float>();
            d_cedge = ctx->cedge.as<float>();
        }
    }
    if ((rc = pinned(ctx->h_seeds, 4ull * (sd + 1) * std::max<uint64_t>(n_seeds_total, 1), "the seed list")))
        return rc;
    if (!noalloc && (rc = ctx->h_ctrl.reserve(sizeof(GhcmCtrl)))) return rc;
    if ((rc = pinned(ctx->rec, sizeof(HostRec), "the progress record"))) return rc;

    uint32_t* hs = ctx->h_seeds.as<uint32_t>();
    {
        size_t k = 0;
        for (uint32_t i = 0; i < n_inst; ++i)
            for (uint32_t s = 0; s < hi[i].n_seeds; ++s) {
                for (uint32_t c = 0; c < sd; ++c) hs[k++] = hi[i].seeds[sd * s + c];
                hs[k++] = i;
            }
    }
    // Grid of the device-mode solve kernel: at most one resident round of work-groups;
    // each work-group loops over list entries, so the grid never has to match the list.
    const uint32_t resident = (uint32_t)std::max(1, sk.blocks_per_cu()) * (uint32_t)std::max(1, ctx->cus);
    const uint32_t grid = std::max(1u, std::min(n_sel, resident));
    uint64_t max_cells = 0;
    for (uint32_t i = 0; i < n_inst; ++i)
        max_cells = std::max<uint64_t>(max_cells, (uint64_t)hi[i].w * hi[i].h * (three_d ? hi[i].d : 1u));
    volatile HostRec* hr = ctx->rec.as<HostRec>();
    hr->progress = 0ull;
    hr->totals_tag = 0u;      // this solve's tag is written by its own k_publish_totals

    SolveArgs a{};
    a.inst = d_inst;
    a.n_inst = n_inst;
    a.tinst = d_tinst;
    a.qwave = d_qwave;
    a.key = d_key;
    a.mask = d_mask;
    a.tmin = d_tmin;
    a.list = d_list;
    a.tiles = T;
    a.shard_cap = shard_cap;
    a.tstat = d_tstat;
    a.jobs = d_sel;
    a.ctrl = d_ctrl;
    a.edge = d_edge;
    a.cedge = d_cedge;
    a.edge_h = edge_h;
    a.eps = cfg.min_change;
    a.iter_cap = cfg.inner_iter_cap;

    // The instance table and the seeds are read by the kernels from mapped pinned memory:
    // the whole solve issues no runtime copy.
    HIP_TRY(hipEventRecord(ctx->ev_begin, st));
    // The time fields at +INF and, for the sweep kernels, the tiles' edge columns in one launch.
    HIP_TRY(launch_solve_init(a, ctx->h_inst.dev_as<GhcmInst>(), n_inst, max_cells, n_sel, d_totals, st,
                              tw, th, td, kind));
    if (three_d)
        HIP_TRY(launch_seed3d(a, ctx->h_seeds.dev_as<uint32_t>(), (uint32_t)n_seeds_total, tw, th, td, st));
    else
        HIP_TRY(launch_seed(a, ctx->h_seeds.dev_as<uint32_t>(), (uint32_t)n_seeds_total, tw, th, st));

    uint64_t max_waves = cfg.max_waves ? cfg.max_waves : (uint64_t)8 * T + 100000;
    // Per-tile counters are summed once after the last wave (outside the timed span) and
    // handed over through the progress record; the host polls for this solve's tag.
    bool totals_ready = false;
    auto publish_totals = [&]() -> int {
        const uint32_t tag = ++ctx->solve_seq;
        HIP_TRY(launch_reduce_stats(a.tstat, T, d_totals, st));
        HIP_TRY(launch_publish_totals(d_totals, a.ctrl, ctx->rec.dev_as<HostRec>(), tag, st));
        for (uint64_t spins = 0; hr->totals_tag != tag; ++spins) {
            if ((spins & 4095u) == 4095u) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady)
                    return fail(GHCM_ERR_HIP, "solve: %s", hipGetErrorString(q));
                if (q == hipSuccess && hr->totals_tag != tag)
                    return fail(GHCM_ERR_HIP, "solve: the progress record was not written");
            }
            __builtin_ia32_pause();
        }
        HIP_TRY(ghcm_capi::spin_wait(ctx->ev_end));
        totals_ready = true;
        return GHCM_OK;
    };
    bool converged = false;
    uint64_t waves_done = 0, visits = 0, inner = 0, improved = 0, kern_launches = 0;
    double kern_us = 0.0;

    // launch_mode 3: tile solve and wave boundary in one launch (sweep kernel).
    const SolveKernel fk = get_fused_kernel((int)tw, (int)th, kind, cfg.kernel);
    const bool fused = !three_d && !host_sched && cfg.launch_mode == 3 && fk.launch != nullptr;
    const PersistentKernel pk = get_persistent_kernel((int)tw, (int)th, kind, cfg.kernel);
    const bool persistent = !three_d && !host_sched && cfg.launch_mode == 1 && pk.launch != nullptr;
    // launch_mode 4: boundary-free waves (sweep kernel): one tile launch per wave; launch_mode 5:
    // the same with the top-n cut made by the tile launch (64 x 64 and 64 x 32 tiles; other
    // shapes run the per-wave launches of launch_mode 0).
    const bool bsel = cfg.launch_mode == 5;
    const BfKernel bk = get_bf_kernel((int)tw, (int)th, kind, cfg.kernel, wide, bsel);
    const bool bfree = !three_d && !host_sched && (cfg.launch_mode == 4 || bsel) && bk.launch != nullptr;
    if (bfree) {
        // ---- boundary-free device wave loop ----
        // Each tile launch reads its own list's counts; its first work-group publishes
        // (done << 32) | (wave + 1) into the mapped progress record.  The host keeps `lead`
        // launches enqueued ahead of the last one started and stops at the first wave that
        // found its list empty; the launches already enqueued behind it exit at once.
        a.rec = ctx->rec.dev_as<HostRec>();
        const bool prof = cfg.profile_kernels != 0;
        const uint32_t lead = std::max<uint32_t>(2, cfg.waves_per_poll ? cfg.waves_per_poll : 4);
        const uint32_t bresident = (uint32_t)std::max(1, bk.blocks_per_cu()) * (uint32_t)std::max(1, ctx->cus);
        const uint32_t bgrid = std::max(1u, std::min(n_sel, bresident));
        uint64_t enqueued = 0;
        for (uint64_t spins = 0;; ++spins) {
            const unsigned long long progress = hr->progress;
            if (progress >> 32) { converged = true; break; }
            if (enqueued >= max_waves) break;
            const uint64_t started = (uint32_t)progress;
            if (enqueued < started + lead) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (prof) {
                    const size_t need = 2 * (enqueued + 1);
                    while (ctx->ev_prof.size() < need) {
                        hipEvent_t ev;
                        HIP_TRY(hipEventCreate(&ev));
                        ctx->ev_prof.push_back(ev);
                    }
                    e0 = ctx->ev_prof[2 * enqueued];
                    e1 = ctx->ev_prof[2 * enqueued + 1];
                }
                HIP_TRY(bk.launch(a, (uint32_t)enqueued, bgrid, st, e0, e1));
                ++enqueued;
                continue;
            }
            if ((spins & 4095u) == 4095u) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady)
                    return fail(GHCM_ERR_HIP, "wave loop: %s", hipGetErrorString(q));
            }
            __builtin_ia32_pause();
        }
        HIP_TRY(hipEventRecord(ctx->ev_end, st));
        if ((rc = publish_totals())) return rc;
        if (prof) {
            for (uint64_t k = 0; k < enqueued; ++k) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_prof[2 * k], ctx->ev_prof[2 * k + 1]));
                kern_us += 1000.0 * ms;
                ++kern_launches;
            }
        }
        converged = (hr->progress >> 32) != 0;
        waves_done = hr->busy_waves;
    } else if (persistent) {
        // ---- one cooperative launch runs every wave (grid = one resident round) ----
        const int per_cu = pk.blocks_per_cu();
        uint32_t pgrid = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, ctx->cus);
        pgrid = std::min<uint32_t>(pgrid, kBarGroup * kMaxBarGroups);
        pgrid = std::max<uint32_t>(1, std::min<uint32_t>(pgrid, std::max<uint32_t>(n_sel, 1)));
        HIP_TRY(launch_wave_end(a, 1u, st));
        if (cfg.profile_kernels) {
            while (ctx->ev_prof.size() < 2) {
                hipEvent_t ev;
                HIP_TRY(hipEventCreate(&ev));
                ctx->ev_prof.push_back(ev);
            }
            HIP_TRY(hipEventRecord(ctx->ev_prof[0], st));
        }
        HIP_TRY(pk.launch(a, pgrid, st));
        if (cfg.profile_kernels) HIP_TRY(hipEventRecord(ctx->ev_prof[1], st));
        HIP_TRY(hipEventRecord(ctx->ev_end, st));
        GhcmCtrl* hc = ctx->h_ctrl.as<GhcmCtrl>();
        HIP_TRY(hipMemcpyAsync(&hc[0], a.ctrl, sizeof(GhcmCtrl), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (cfg.profile_kernels) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_prof[0], ctx->ev_prof[1]));
            kern_us += 1000.0 * ms;
            kern_launches += 1;
        }
        if (hc[0].abort) return fail(GHCM_ERR_HIP, "device wave loop aborted: a grid wait timed out");
        converged = hc[0].done != 0;
        waves_done = hc[0].busy_waves;
    } else if (!host_sched) {
        // ---- device-resident wave loop ----
        // k_wave_end publishes a tick, the done flag and the busy-wave count into the
        // mapped progress record on every launch.  The host keeps `lead` waves queued
        // ahead of the last executed boundary and stops as soon as it reads `done`, so at
        // most `lead` idle waves (two cheap early-exit launches each) follow the last
        // busy one, and no runtime copy or blocking wait sits in the loop.
        a.rec = ctx->rec.dev_as<HostRec>();
        HIP_TRY(launch_wave_end(a, 1u, st));
        const bool prof = cfg.profile_kernels != 0;
        const uint32_t lead = std::max<uint32_t>(2, cfg.waves_per_poll ? cfg.waves_per_poll : 4);
        const bool use_graph = cfg.launch_mode == 2 && !prof;
        // The fused kernel's grid: one resident round of its (larger-LDS) work-groups.
        const uint32_t fgrid = fused ? std::max(1u, std::min(n_sel, (uint32_t)std::max(1, fk.blocks_per_cu()) *
                                                                      (uint32_t)std::max(1, ctx->cus)))
                                     : 0u;
        const uint32_t step = use_graph ? std::max<uint32_t>(1, lead / 2) : 1;
        if (use_graph) {
            // key: the kernel (tile shape, field kind and style are one launcher each), its
            // arguments and the launch shape of one captured step of waves
            std::vector<unsigned char> key(sizeof(SolveArgs) + 2 * sizeof(uint32_t) + sizeof(launch));
            unsigned char* kp = key.data();
            std::memcpy(kp, &a, sizeof(SolveArgs));
            std::memcpy(kp + sizeof(SolveArgs), &grid, sizeof(uint32_t));
            std::memcpy(kp + sizeof(SolveArgs) + sizeof(uint32_t), &step, sizeof(uint32_t));
            std::memcpy(kp + sizeof(SolveArgs) + 2 * sizeof(uint32_t), &launch, sizeof(launch));
            if (!ctx->gexec || key != ctx->gkey) {
                if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
                ctx->gexec = nullptr;
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipError_t ce = hipSuccess;
                for (uint32_t k = 0; k < step && ce == hipSuccess; ++k) {
                    ce = launch(a, grid, st, nullptr, nullptr);
                    if (ce == hipSuccess) ce = launch_wave_end(a, 0u, st);
                }
                hipError_t ee = hipStreamEndCapture(st, &graph);
                if (ce != hipSuccess || ee != hipSuccess)
                    return fail(GHCM_ERR_HIP, "graph capture failed: %s",
                                hipGetErrorString(ce != hipSuccess ? ce : ee));
                hipError_t ie = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) {
                    ctx->gexec = nullptr;
                    return fail(GHCM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
                }
                ctx->gkey = key;
            }
        }
        uint64_t enqueued = 0;                      // waves queued after the first boundary
        for (uint64_t spins = 0;; ++spins) {
            const unsigned long long progress = hr->progress;
            if (progress >> 32) { converged = true; break; }
            if (enqueued >= max_waves) break;
            const uint32_t ticks = (uint32_t)progress;
            const uint64_t executed = ticks ? ticks - 1u : 0u;
            if (enqueued < executed + lead) {
                if (use_graph) {
                    HIP_TRY(hipGraphLaunch(ctx->gexec, st));
                } else if (prof) {
                    const size_t need = 2 * (enqueued + 1);
                    while (ctx->ev_prof.size() < need) {
                        hipEvent_t ev;
                        HIP_TRY(hipEventCreate(&ev));
                        ctx->ev_prof.push_back(ev);
                    }
                    hipEvent_t e0 = ctx->ev_prof[2 * enqueued], e1 = ctx->ev_prof[2 * enqueued + 1];
                    if (fused) {
                        HIP_TRY(fk.launch(a, fgrid, st, e0, e1));
                    } else {
                        HIP_TRY(launch(a, grid, st, e0, e1));
                        HIP_TRY(launch_wave_end(a, 0u, st));
                    }
                } else if (fused) {
                    HIP_TRY(fk.launch(a, fgrid, st, nullptr, nullptr));
                } else {
                    HIP_TRY(launch(a, grid, st, nullptr, nullptr));
                    HIP_TRY(launch_wave_end(a, 0u, st));
                }
                enqueued += step;
                continue;
            }
            if ((spins & 4095u) == 4095u) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady)
                    return fail(GHCM_ERR_HIP, "wave loop: %s", hipGetErrorString(q));
            }
            __builtin_ia32_pause();
        }
        HIP_TRY(hipEventRecord(ctx->ev_end, st));
        if ((rc = publish_totals())) return rc;
        if (prof) {
            for (uint64_t k = 0; k < enqueued; ++k) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_prof[2 * k], ctx->ev_prof[2 * k + 1]));
                kern_us += 1000.0 * ms;
                ++kern_launches;
            }
        }
        converged = (hr->progress >> 32) != 0;
        waves_done = hr->busy_waves;
    } else {
        // ---- host priority queue (src/eikonal_hcm.cpp:343-482) ----
        if ((rc = ctx->jobs.reserve(sizeof(HostJob) * n_sel))) return rc;
        if ((rc = ctx->status.reserve(sizeof(HostStatus) * n_sel))) return rc;
        if ((rc = ctx->h_jobs.reserve(sizeof(HostJob) * n_sel))) return rc;
        if ((rc = ctx->h_status.reserve(sizeof(HostStatus) * n_sel))) return rc;
        a.host_jobs = ctx->jobs.as<HostJob>();
        a.status = ctx->status.as<HostStatus>();
        HostJob* hj = ctx->h_jobs.as<HostJob>();
        HostStatus* hst = ctx->h_status.as<HostStatus>();
        std::vector<float> smallest(T, INFINITY);
        std::vector<uint32_t> border(T, 0);
        std::vector<uint8_t> dirty(T, 0);
        std::vector<uint32_t> inst_of(T);
        for (uint32_t i = 0; i < n_inst; ++i)
            std::fill(inst_of.begin() + hinst[i].tile_base,
                      inst_of.begin() + hinst[i].tile_base + hinst[i].tiles_x * hinst[i].tiles_y, i);
        auto nbr = [&](uint32_t id, int s, uint32_t* out) -> bool {
            const GhcmInst& g = hinst[inst_of[id]];
            int tx, ty;
            tile_xy(id - g.tile_base, g.tiles_x, tx, ty);
            *out = tile_neighbour(g.tile_base, g.tiles_x, g.tiles_y, tx, ty, s);
            return *out != kNoTile;
        };
        // BlockWrapper key: min smallest_time of the tile and its 4 neighbours, frozen
        // at push (src/eikonal_hcm.cpp:47-80).
        auto key_of = [&](uint32_t id) {
            float k = smallest[id];
            uint32_t nb;
            for (int s = 0; s < 4; ++s)
                if (nbr(id, s, &nb)) k = std::min(k, smallest[nb]);
            return k;
        };
        struct Ent {
            float key;
            uint32_t id;
            bool operator<(const Ent& o) const { return key > o.key; }
        };
        std::priority_queue<Ent> pq;
        while (cfg.profile_kernels && ctx->ev_prof.size() < 2) {
            hipEvent_t ev;
            HIP_TRY(hipEventCreate(&ev));
            ctx->ev_prof.push_back(ev);
        }
        for (uint32_t i = 0; i < n_inst; ++i)
            for (uint32_t s = 0; s < hi[i].n_seeds; ++s) {
                const int dx[5] = {0, -1, 1, 0, 0}, dy[5] = {0, 0, 0, -1, 1};
                for (int d = 0; d < 5; ++d) {
                    int px = (int)hi[i].seeds[2 * s] + dx[d], py = (int)hi[i].seeds[2 * s + 1] + dy[d];
                    if (px < 0 || py < 0 || px >= (int)hi[i].w || py >= (int)hi[i].h) continue;
                    const uint32_t t = tile_of_cell(hinst[i].tile_base, hinst[i].tiles_x, (uint32_t)px,
                                                    (uint32_t)py, tw, th);
                    border[t] |= kMaskOpenAll;
                    smallest[t] = 0.0f;
                    if (!dirty[t]) { dirty[t] = 1; pq.push({0.0f, t}); }
                }
            }
        // launch_mode 2: a wave is one replay of a graph captured once -- the upload of n job
        // slots, the tile launch over them and the read-back of their statuses -- instead of
        // three enqueued operations.  Slots past the wave's jobs hold kNoTile, which the
        // kernels skip.
        const bool pq_graph = cfg.launch_mode == 2 && !cfg.profile_kernels;
        uint32_t filled = 0;                    // slots [0, filled) may hold a stale tile id
        if (pq_graph) {
            for (uint32_t j = 0; j < n_sel; ++j) hj[j] = HostJob{kNoTile, 0u};
            SolveArgs ga = a;
            ga.n_host_jobs = n_sel;
            std::vector<unsigned char> key(sizeof(SolveArgs) + sizeof(launch) + 2 * sizeof(void*) + 1);
            unsigned char* kp = key.data();
            std::memcpy(kp, &ga, sizeof(SolveArgs));
            std::memcpy(kp + sizeof(SolveArgs), &launch, sizeof(launch));
            std::memcpy(kp + sizeof(SolveArgs) + sizeof(launch), &hj, sizeof(void*));
            std::memcpy(kp + sizeof(SolveArgs) + sizeof(launch) + sizeof(void*), &hst, sizeof(void*));
            kp[key.size() - 1] = 0x9A;          // tells host-queue graphs from device-loop ones
            if (!ctx->gexec || key != ctx->gkey) {
                if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
                ctx->gexec = nullptr;
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipError_t ce = hipMemcpyAsync(ctx->jobs.p, hj, sizeof(HostJob) * n_sel, hipMemcpyHostToDevice, st);
                if (ce == hipSuccess) ce = launch(ga, n_sel, st, nullptr, nullptr);
                if (ce == hipSuccess)
                    ce = hipMemcpyAsync(hst, ga.status, sizeof(HostStatus) * n_sel, hipMemcpyDeviceToHost, st);
                hipError_t ee = hipStreamEndCapture(st, &graph);
                if (ce != hipSuccess || ee != hipSuccess)
                    return fail(GHCM_ERR_HIP, "graph capture failed: %s",
                                hipGetErrorString(ce != hipSuccess ? ce : ee));
                hipError_t ie = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) {
                    ctx->gexec = nullptr;
                    return fail(GHCM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
                }
                ctx->gkey = key;
            }
        }
        while (!pq.empty()) {
            if (waves_done >= max_waves) break;
            uint32_t nj = 0;
            while (nj < n_sel && !pq.empty()) {
                uint32_t id = pq.top().id;
                pq.pop();
                hj[nj].id = id;
                hj[nj].border = border[id];
                border[id] = 0;
                dirty[id] = 0;   // cleared before results, so same-wave neighbours re-queue
                ++nj;
            }
            if (pq_graph) {
                for (uint32_t j = nj; j < filled; ++j) hj[j] = HostJob{kNoTile, 0u};
                filled = nj;
                HIP_TRY(hipGraphLaunch(ctx->gexec, st));
            } else {
                a.n_host_jobs = nj;
                HIP_TRY(hipMemcpyAsync(ctx->jobs.p, hj, sizeof(HostJob) * nj, hipMemcpyHostToDevice, st));
                if (cfg.profile_kernels)
                    HIP_TRY(launch(a, nj, st, ctx->ev_prof[0], ctx->ev_prof[1]));
                else
                    HIP_TRY(launch(a, nj, st, nullptr, nullptr));
                HIP_TRY(hipMemcpyAsync(hst, a.status, sizeof(HostStatus) * nj, hipMemcpyDeviceToHost, st));
            }
            HIP_TRY(ghcm_capi::spin_sync(st, ctx->ev_sync));
            if (cfg.profile_kernels) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_prof[0], ctx->ev_prof[1]));
                kern_us += 1000.0 * ms;
                ++kern_launches;
            }
            ++waves_done;
            for (uint32_t j = 0; j < nj; ++j) {
                const uint32_t id = hj[j].id;
                const HostStatus& s = hst[j];
                smallest[id] = s.smallest_time;
                for (int side = 0; side < 4; ++side) {
                    uint32_t nb;
                    if (!s.border[side] || !nbr(id, side, &nb)) continue;
                    border[nb] |= border_bit_across(side);
                    if (!dirty[nb]) { dirty[nb] = 1; pq.push({key_of(nb), nb}); }
                }
                if (s.requeue) {
                    border[id] |= kMaskOpenAll;
                    if (!dirty[id]) { dirty[id] = 1; pq.push({key_of(id), id}); }
                }
            }
        }
        converged = pq.empty();
        HIP_TRY(hipEventRecord(ctx->ev_end, st));
    }

    if (!totals_ready && (rc = publish_totals())) return rc;
    visits = hr->totals[0];
    inner = hr->totals[1];
    improved = hr->totals[2];
    float dev_ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&dev_ms, ctx->ev_begin, ctx->ev_end));
    const auto t_wall1 = std::chrono::steady_clock::now();
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->duration_total_us = std::chrono::duration<double, std::micro>(t_wall1 - t_wall0).count();
        stats->duration_device_us = 1000.0 * dev_ms;
        stats->waves = waves_done;
        stats->tile_visits = visits;
        stats->tiles_total = T;
        stats->cells = cells;
        stats->inner_iterations = inner;
        stats->cells_improved = improved;
        stats->converged = converged ? 1u : 0u;
        stats->tile_w = tw;
        stats->tile_h = th;
        stats->tile_d = td;
        stats->n_active = n_sel;
        stats->bytes_per_visit = three_d ? bytes_per_visit3d(tw, th, td, kind) : bytes_per_visit(tw, th, kind);
        stats->solve_kernel_us = kern_us;
        stats->solve_launches = kern_launches;
    }
    if (!converged) return fail(GHCM_ERR_NOT_CONVERGED, "wave cap %llu reached", (unsigned long long)max_waves);
    return GHCM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// solve entry points
// ---------------------------------------------------------------------------------------

// Host-buffer copies go through two 16 MB pinned parts: the host copy of one part (split over
// threads) overlaps the DMA of the other.  A pageable hipMemcpyAsync of the 4096^2 field and
// times moved about 14 GB/s.
constexpr size_t kIoPart = size_t(16) << 20;

int io_prepare(ghcm_ctx* ctx) {
    for (int b = 0; b < 2; ++b) {
        if (int rc = ctx->io_stage[b].reserve(kIoPart)) return rc;
        if (!ctx->ev_io[b]) HIP_TRY(hipEventCreateWithFlags(&ctx->ev_io[b], hipEventDisableTiming));
    }
    return GHCM_OK;
}

int stage_in(ghcm_ctx* ctx, void* d_dst, const void* h_src, size_t n) {
    if (int rc = io_prepare(ctx)) return rc;
    int b = 0;
    for (size_t off = 0; off < n; off += kIoPart, b ^= 1) {
        const size_t len = std::min(kIoPart, n - off);
        HIP_TRY(ghcm_capi::spin_wait(ctx->ev_io[b]));       // part b's previous DMA is done
        ghcm_capi::par_memcpy(ctx->io_stage[b].p, static_cast<const char*>(h_src) + off, len);
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(d_dst) + off, ctx->io_stage[b].p, len,
                               hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev_io[b], ctx->stream));
    }
    return GHCM_OK;
}

int stage_out(ghcm_ctx* ctx, void* h_dst, const void* d_src, size_t n) {
    if (int rc = io_prepare(ctx)) return rc;
    const size_t parts = (n + kIoPart - 1) / kIoPart;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t off = i * kIoPart, len = std::min(kIoPart, n - off);
        hipError_t e = hipMemcpyAsync(ctx->io_stage[i & 1].p, static_cast<const char*>(d_src) + off,
                                      len, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_io[i & 1], ctx->stream);
        return e;
    };
    for (size_t i = 0; i < parts && i < 2; ++i) HIP_TRY(issue(i));
    for (size_t i = 0; i < parts; ++i) {
        const size_t off = i * kIoPart, len = std::min(kIoPart, n - off);
        HIP_TRY(ghcm_capi::spin_wait(ctx->ev_io[i & 1]));
        ghcm_capi::par_memcpy(static_cast<char*>(h_dst) + off, ctx->io_stage[i & 1].p, len);
        if (i + 2 < parts) HIP_TRY(issue(i + 2));
    }
    return GHCM_OK;
}

extern "C" int ghcm_solve_device(ghcm_ctx* ctx, const void* d_field, int field_kind, uint32_t w,
                                 uint32_t h, const uint32_t* seeds_xy, uint32_t n_seeds,
                                 float* d_time, ghcm_stats* stats) {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    HostInst hi{d_field, d_time, w, h, seeds_xy, n_seeds};
    return solve_core(ctx, field_kind, &hi, 1, stats);
}

extern "C" int ghcm_solve(ghcm_ctx* ctx, const void* field, int field_kind, uint32_t w, uint32_t h,
                          const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
                          ghcm_stats* stats) {
  // The staging copies allocate (the copy pool's queue): nothing may escape the C boundary.
  return no_throw([&]() -> int {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    if (!field || !time_out) return fail(GHCM_ERR_INVALID_ARG, "null field or output");
    if (!valid_kind(field_kind)) return fail(GHCM_ERR_INVALID_ARG, "unknown field kind %d", field_kind);
    if (w == 0 || h == 0) return fail(GHCM_ERR_INVALID_ARG, "zero-sized field");
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(ctx->ordinal));
    const size_t n = (size_t)w * h;
    int rc = ctx->field_io.reserve(n * kind_bytes(field_kind));
    if (rc) return rc;
    if ((rc = ctx->time_io.reserve(n * 4))) return rc;
    if ((rc = stage_in(ctx, ctx->field_io.p, field, n * kind_bytes(field_kind)))) return rc;
    HostInst hi{ctx->field_io.p, ctx->time_io.as<float>(), w, h, seeds_xy, n_seeds};
    rc = solve_core(ctx, field_kind, &hi, 1, stats);
    if (rc != GHCM_OK && rc != GHCM_ERR_NOT_CONVERGED) return rc;
    if (int rc2 = stage_out(ctx, time_out, ctx->time_io.p, n * 4)) return rc2;
    if (stats)
        stats->duration_total_us =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  });
}

extern "C" int ghcm_solve_device_noalloc(ghcm_ctx* ctx, const void* d_field, int field_kind, uint32_t w,
                                         uint32_t h, const uint32_t* seeds_xy, uint32_t n_seeds,
                                         float* d_time, ghcm_stats* stats) {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    if (n_seeds > kNoAllocSeeds)
        return fail(GHCM_ERR_INVALID_ARG, "allocation-free solve: %u seeds, at most %u", n_seeds, kNoAllocSeeds);
    HostInst hi{d_field, d_time, w, h, seeds_xy, n_seeds};
    return solve_core(ctx, field_kind, &hi, 1, stats, false, true);
}

extern "C" int ghcm_solve3d_device(ghcm_ctx* ctx, const void* d_field, int field_kind, uint32_t w,
                                   uint32_t h, uint32_t d, const uint32_t* seeds_xyz, uint32_t n_seeds,
                                   float* d_time, ghcm_stats* stats) {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    HostInst hi{d_field, d_time, w, h, seeds_xyz, n_seeds, d};
    return solve_core(ctx, field_kind, &hi, 1, stats, true);
}

extern "C" int ghcm_solve3d(ghcm_ctx* ctx, const void* field, int field_kind, uint32_t w, uint32_t h,
                            uint32_t d, const uint32_t* seeds_xyz, uint32_t n_seeds, float* time_out,
                            ghcm_stats* stats) {
  return no_throw([&]() -> int {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    if (!field || !time_out) return fail(GHCM_ERR_INVALID_ARG, "null field or output");
    if (!valid_kind(field_kind)) return fail(GHCM_ERR_INVALID_ARG, "unknown field kind %d", field_kind);
    if (w == 0 || h == 0 || d == 0) return fail(GHCM_ERR_INVALID_ARG, "zero-sized field");
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(ctx->ordinal));
    const size_t n = (size_t)w * h * d;
    int rc = ctx->field_io.reserve(n * kind_bytes(field_kind));
    if (rc) return rc;
    if ((rc = ctx->time_io.reserve(n * 4))) return rc;
    if ((rc = stage_in(ctx, ctx->field_io.p, field, n * kind_bytes(field_kind)))) return rc;
    HostInst hi{ctx->field_io.p, ctx->time_io.as<float>(), w, h, seeds_xyz, n_seeds, d};
    rc = solve_core(ctx, field_kind, &hi, 1, stats, true);
    if (rc != GHCM_OK && rc != GHCM_ERR_NOT_CONVERGED) return rc;
    if (int rc2 = stage_out(ctx, time_out, ctx->time_io.p, n * 4)) return rc2;
    if (stats)
        stats->duration_total_us =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  });
}

extern "C" int ghcm_dev_alloc(ghcm_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->ordinal));
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(GHCM_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return GHCM_OK;
}

extern "C" int ghcm_dev_free(ghcm_ctx* ctx, void* p) {
    if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
    if (!p) return GHCM_OK;
    HIP_TRY(hipSetDevice(ctx->ordinal));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(p));
    return GHCM_OK;
}

extern "C" int ghcm_dev_upload(ghcm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->ordinal));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GHCM_OK;
}

extern "C" int ghcm_dev_download(ghcm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->ordinal));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GHCM_OK;
}

extern "C" int ghcm_dev_copy(ghcm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->ordinal));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GHCM_OK;
}

extern "C" int ghcm_rgb_luminosity_device(ghcm_ctx* ctx, const float* d_rgb, uint32_t w, uint32_t h,
                                          float* d_out) {
    if (!ctx || !d_rgb || !d_out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    if (!w || !h) return fail(GHCM_ERR_INVALID_ARG, "zero-sized image");
    HIP_TRY(hipSetDevice(ctx->ordinal));
    HIP_TRY(launch_rgb_luminosity(d_rgb, d_out, (size_t)w * h, ctx->stream));
    HIP_TRY(ghcm_capi::spin_sync(ctx->stream, ctx->ev_sync));
    return GHCM_OK;
}

extern "C" int ghcm_solve_batch(ghcm_ctx* ctx, int field_kind, const ghcm_instance* inst,
                                uint32_t n_inst, ghcm_stats* stats) {
    return no_throw([&]() -> int {
        if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
        if (!inst || n_inst == 0) return fail(GHCM_ERR_INVALID_ARG, "no instances");
        std::vector<HostInst> hi(n_inst);
        for (uint32_t i = 0; i < n_inst; ++i)
            hi[i] = HostInst{inst[i].d_field, inst[i].d_time, inst[i].w, inst[i].h, inst[i].seeds_xy, inst[i].n_seeds};
        return solve_core(ctx, field_kind, hi.data(), n_inst, stats);
    });
}

// ---------------------------------------------------------------------------------------
// test support
// ---------------------------------------------------------------------------------------

extern "C" int ghcm_tile_info(uint32_t w, uint32_t h, uint32_t tile_w, uint32_t tile_h,
                              uint32_t tile_base, uint32_t tile, ghcm_tile_desc* out) {
    if (!out) return fail(GHCM_ERR_INVALID_ARG, "out is null");
    if (!w || !h || !tile_w || !tile_h) return fail(GHCM_ERR_INVALID_ARG, "zero size");
    const uint32_t tx_n = tiles_along(w, tile_w), ty_n = tiles_along(h, tile_h);
    if ((uint64_t)tile_base + (uint64_t)tx_n * ty_n > 0xFFFFFFFFull)
        return fail(GHCM_ERR_INVALID_ARG, "tile ids overflow");
    if (tile < tile_base || tile - tile_base >= tx_n * ty_n)
        return fail(GHCM_ERR_INVALID_ARG, "tile %u outside [%u, %u)", tile, tile_base, tile_base + tx_n * ty_n);
    std::memset(out, 0, sizeof *out);
    int tx, ty;
    tile_xy(tile - tile_base, tx_n, tx, ty);
    out->tiles_x = tx_n;
    out->tiles_y = ty_n;
    out->tx = (uint32_t)tx;
    out->ty = (uint32_t)ty;
    out->x0 = (uint32_t)tx * tile_w;
    out->y0 = (uint32_t)ty * tile_h;
    out->w = (uint32_t)tile_extent((int)w, (int)out->x0, (int)tile_w);
    out->h = (uint32_t)tile_extent((int)h, (int)out->y0, (int)tile_h);
    for (int s = 0; s < 4; ++s) {
        out->neighbour[s] = tile_neighbour(tile_base, tx_n, ty_n, tx, ty, s);
        out->border_bit[s] = border_bit_across(s);
    }
    out->seed_tile = tile_of_cell(tile_base, tx_n, out->x0 + out->w - 1, out->y0 + out->h - 1, tile_w, tile_h);
    return GHCM_OK;
}

extern "C" int ghcm_debug_select(ghcm_ctx* ctx, uint32_t T, const uint32_t* tiles, const float* keys,
                                 const uint32_t* masks, uint32_t count, uint32_t n, int handoff,
                                 uint32_t* jobs, uint32_t* n_jobs, uint32_t* handed, uint32_t* n_handed,
                                 uint32_t* next, float* next_keys, uint32_t* next_masks,
                                 uint32_t* n_next) {
    return no_throw([&]() -> int {
        if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
        if (!T || count > T || !n) return fail(GHCM_ERR_INVALID_ARG, "need 0 < n, count <= tiles_total");
        if (count && (!tiles || !keys || !masks || !jobs || !handed || !next || !next_keys || !next_masks))
            return fail(GHCM_ERR_INVALID_ARG, "null buffer");
        if (!n_jobs || !n_handed || !n_next) return fail(GHCM_ERR_INVALID_ARG, "null count");
        std::vector<uint8_t> seen(T, 0);
        for (uint32_t i = 0; i < count; ++i) {
            if (tiles[i] >= T || seen[tiles[i]]) return fail(GHCM_ERR_INVALID_ARG, "tile %u out of range or repeated", tiles[i]);
            seen[tiles[i]] = 1;
            if (!(keys[i] >= 0.0f)) return fail(GHCM_ERR_INVALID_ARG, "key %u is negative or NaN", i);
        }
        HIP_TRY(hipSetDevice(ctx->ordinal));
        hipStream_t st = ctx->stream;
        const uint32_t shard_cap = (T + kShards - 1) / kShards;
        int rc;
        if ((rc = ctx->qwave.reserve(4ull * T))) return rc;
        if ((rc = ctx->key.reserve(8ull * T))) return rc;
        if ((rc = ctx->mask.reserve(8ull * T))) return rc;
        if ((rc = ctx->list.reserve(12ull * kShards * shard_cap))) return rc;
        if ((rc = ctx->ctrl.reserve(sizeof(GhcmCtrl)))) return rc;
        if ((rc = ctx->sel.reserve(4ull * T))) return rc;
        // Wave W's list is list 0.  The first boundary selects from it as it stands; a later
        // boundary first swaps lists (it empties list 1, the one its wave consumed).
        const uint32_t W = handoff ? 1u : 0u, p = W & 1u;
        std::vector<uint32_t> hkey(2ull * T, kInfBits), hmask(2ull * T, 0u), hq(T, 0u);
        std::vector<uint32_t> hlist(3ull * kShards * shard_cap, 0u);
        std::vector<uint32_t> fill(kShards, 0u);
        GhcmCtrl hc;
        std::memset(&hc, 0, sizeof hc);
        for (uint32_t i = 0; i < count; ++i) {
            const uint32_t g = tiles[i], sh = g % kShards;
            hlist[(size_t)sh * shard_cap + fill[sh]++] = g;
            uint32_t kb;
            std::memcpy(&kb, &keys[i], 4);
            hkey[(size_t)p * T + g] = kb;
            hmask[(size_t)p * T + g] = masks[i];
            hq[g] = W + 1u;
        }
        for (uint32_t sh = 0; sh < kShards; ++sh) hc.count[0][sh * kCounterStride] = fill[sh];
        hc.cur = handoff ? 1u : 0u;
        hc.wave = 0u;
        hc.n_select = n;
        HIP_TRY(hipMemcpyAsync(ctx->ctrl.p, &hc, sizeof hc, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ctx->key.p, hkey.data(), 8ull * T, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ctx->mask.p, hmask.data(), 8ull * T, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ctx->qwave.p, hq.data(), 4ull * T, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(ctx->list.p, hlist.data(), hlist.size() * 4, hipMemcpyHostToDevice, st));
        SolveArgs a{};
        a.qwave = ctx->qwave.as<uint32_t>();
        a.key = ctx->key.as<uint32_t>();
        a.mask = ctx->mask.as<uint32_t>();
        a.list = ctx->list.as<uint32_t>();
        a.tiles = T;
        a.shard_cap = shard_cap;
        a.jobs = ctx->sel.as<uint32_t>();
        a.ctrl = ctx->ctrl.as<GhcmCtrl>();
        HIP_TRY(launch_wave_end(a, handoff ? 0u : 1u, st));
        std::vector<uint32_t> hsel(T);
        HIP_TRY(hipMemcpyAsync(&hc, ctx->ctrl.p, sizeof hc, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(hsel.data(), ctx->sel.p, 4ull * T, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(hlist.data(), ctx->list.p, hlist.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(hkey.data(), ctx->key.p, 8ull * T, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(hmask.data(), ctx->mask.p, 8ull * T, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const uint32_t cur = hc.cur;          // the list the boundary selected from
        if (hc.n_jobs > count || hc.n_carry > count - std::min(count, hc.n_jobs))
            return fail(GHCM_ERR_HIP, "selection counts out of range (%u jobs, %u carried)", hc.n_jobs, hc.n_carry);
        *n_jobs = hc.n_jobs;
        if (hc.jobs_from_list) {
            uint32_t k = 0;
            for (uint32_t sh = 0; sh < kShards && k < hc.n_jobs; ++sh)
                for (uint32_t i = 0; i < hc.count[cur][sh * kCounterStride] && k < hc.n_jobs; ++i)
                    jobs[k++] = hlist[((size_t)cur * kShards + sh) * shard_cap + i];
        } else {
            for (uint32_t k = 0; k < hc.n_jobs; ++k) jobs[k] = hsel[k];
        }
        *n_handed = hc.n_carry;
        for (uint32_t k = 0; k < hc.n_carry; ++k) handed[k] = hsel[hc.n_jobs + k];
        const uint32_t nl = cur ^ 1u, np = (W + 1u) & 1u;
        uint32_t m = 0;
        for (uint32_t sh = 0; sh < kShards; ++sh) {
            const uint32_t c = hc.count[nl][sh * kCounterStride];
            if (m + c > count || c > shard_cap) return fail(GHCM_ERR_HIP, "next list longer than the input");
            for (uint32_t i = 0; i < c; ++i, ++m) {
                const uint32_t g = hlist[((size_t)nl * kShards + sh) * shard_cap + i];
                if (g >= T) return fail(GHCM_ERR_HIP, "next list entry %u out of range", g);
                next[m] = g;
                std::memcpy(&next_keys[m], &hkey[(size_t)np * T + g], 4);
                next_masks[m] = hmask[(size_t)np * T + g];
            }
        }
        *n_next = m;
        return (int)GHCM_OK;
    });
}
