// serve.cpp -- continuous-batching solve server (ghcm_serve_*).
//
// One device wave loop runs for the lifetime of the server and never restarts between
// requests.  A request is admitted between waves when an instance slot, a range of tile
// state and (for host buffers) room for its field and time buffer inside the HBM arena
// are all free, first come first served.  Admission is a few launches on the loop's
// stream: the slot's instance entry, admission generation and tile range are reset
// (k_admit_tiles), the time buffer is set to +INF, the seeds queue their tiles for the
// next wave (k_seed_serve, as a plain solve's seeds are queued for its first wave,
// src/eikonal_hcm.cpp:189-238) and the loop leaves its idle state (k_wake).  From then
// on the request's tiles compete in the same top-n selection as every live request's.
//
// Every wave boundary (k_wave_end) writes, into mapped pinned host memory, each slot's
// queued-tile count tagged with the slot's generation.  A request is finished at the
// first boundary that reports its generation with a count of 0: tiles are only queued
// by seeds or by a solved neighbour of the same instance, so nothing can reopen it.
// Host-buffer requests are staged through pinned blocks and moved by copy kernels on two
// side streams (upload before admission, download after completion), ordered against the
// loop by events; the loop itself never waits on the runtime.
//
// The reference solves one field per call (src/eikonal_hcm.cpp:186-529); this loop is
// the serving form of the same wave algorithm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "capi_common.h"
#include "ghcm.h"
#include "ghcm_dev.h"
#include "kernels.h"

using namespace ghcm;
using ghcm_capi::DevBuf;
using ghcm_capi::PinnedBuf;
using ghcm_capi::par_memcpy;
using ghcm_capi::fail;
using ghcm_capi::kind_bytes;
using ghcm_capi::no_throw;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) with a coalescing free list.
class RangeAlloc {
public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, next = it->first + n;
            free_.erase(it);
            if (rest) free_[next] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        if (!n) return;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0;
};

// Pinned host blocks for staging request fields and results, reused across requests.
// Sizes are rounded up to a quarter-power-of-two class so blocks fit later requests.
class PinnedPool {
public:
    static size_t size_class(size_t b) {
        size_t p = 1u << 16;
        while (p < b && p * 2 <= b) p <<= 1;
        if (p >= b) return p;
        for (int q = 5; q <= 8; ++q)
            if (p / 4 * q >= b) return p / 4 * q;
        return p * 2;
    }
    // A block of at least `bytes`, mapped for kernels (*dev: its device address).
    void* get(size_t bytes, void** dev) {
        const size_t c = size_class(bytes);
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.find(c);
            if (it != free_.end() && !it->second.empty()) {
                void* p = it->second.back();
                it->second.pop_back();
                *dev = dev_[p];
                return p;
            }
        }
        void* p = nullptr;
        void* d = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
            (void)hipHostFree(p);
            return nullptr;
        }
        std::lock_guard<std::mutex> lk(mu_);
        all_.push_back(p);
        dev_[p] = d;
        bytes_ += c;
        *dev = d;
        return p;
    }
    void put(void* p, size_t bytes) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_[size_class(bytes)].push_back(p);
    }
    void release() {
        std::lock_guard<std::mutex> lk(mu_);
        for (void* p : all_) (void)hipHostFree(p);
        all_.clear();
        free_.clear();
        dev_.clear();
        bytes_ = 0;
    }
    uint64_t bytes() {
        std::lock_guard<std::mutex> lk(mu_);
        return bytes_;
    }

private:
    std::mutex mu_;
    std::map<size_t, std::vector<void*>> free_;
    std::vector<void*> all_;
    std::unordered_map<void*, void*> dev_;
    uint64_t bytes_ = 0;
};

constexpr uint64_t kAlign = 256;
inline uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

enum class ReqState { Pending, Live, Copying, Done };

struct Request {
    uint64_t ticket = 0;
    bool device = false;                // device-resident request (no staging, no arena)
    const void* d_field = nullptr;      // device request: caller's buffers
    float* d_time = nullptr;
    void* h_field = nullptr;            // host request: pinned staging of the field ...
    void* h_field_dev = nullptr;
    void* h_time = nullptr;             // ... and of the result
    void* h_time_dev = nullptr;
    uint32_t w = 0, h = 0;
    std::vector<uint32_t> seeds;        // (x, y) pairs
    float* time_out = nullptr;          // host request: caller's result buffer
    bool waited = false;                // a ghcm_serve_wait is collecting it
    ReqState state = ReqState::Pending;
    int status = GHCM_OK;
    Clock::time_point t_submit, t_admit, t_done;
    // placement
    uint32_t slot = 0, gen = 0;
    uint64_t arena_off = 0, arena_bytes = 0;
    uint64_t tile_base = 0, tiles = 0;
    uint32_t tiles_x = 0, tiles_y = 0;
    uint32_t admit_tick = 0;            // tick of the first boundary after admission
    uint64_t waves = 0;
    hipEvent_t ev_ready = nullptr;      // host request: the loop passed its completion
    hipEvent_t ev_done = nullptr;       // host request: result copy landed
};

}  // namespace

struct ghcm_server {
    int ordinal = 0;
    ghcm_config cfg{};
    ghcm_serve_config scfg{};
    int kind = GHCM_FIELD_COST_F32;
    uint32_t tw = 0, th = 0, n_sel = 0, grid = 1, lead = 8, step = 1, slots = 0, max_tiles = 0;
    uint64_t budget = 0;
    solve_launcher_t launch = nullptr;
    hipStream_t stream = nullptr, up_stream = nullptr, down_stream = nullptr;
    hipEvent_t ev_upload = nullptr;
    hipGraphExec_t gexec = nullptr;
    DevBuf arena, inst, tinst, qwave, key, mask, tmin, list, tstat, sel, ctrl, sgen, totals;
    DevBuf edge, cedge;                  // sweep kernel: edge columns of every tile of the pool
    static constexpr int kSeedRing = 8;   // seed staging buffers, reused round robin
    PinnedBuf h_seeds[kSeedRing], rec;
    hipEvent_t ev_seed[kSeedRing] = {};  // recorded after the launch that read each buffer
    int seed_next = 0;
    PinnedPool pool;
    SolveArgs a{};
    RangeAlloc arena_alloc, tile_alloc;
    std::vector<uint32_t> free_slots;
    std::vector<Request*> live;          // by slot (server thread only)
    std::vector<hipEvent_t> ev_pool;     // server thread only
    uint32_t next_gen = 1;               // server thread only

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> tickets;
    uint64_t next_ticket = 1;
    bool stop = false, drain = true, exited = false;
    uint32_t waiters = 0;                // ghcm_serve_wait calls still inside (under mu)
    int fatal = GHCM_OK;
    std::string fatal_msg;
    ghcm_serve_stats st{};
    std::thread thread;

    int init();
    void run();
    int loop();
    int admit(std::vector<Request*>& batch, uint32_t enqueued);
    int enqueue_step();
    int get_event(hipEvent_t* ev);
    bool place(Request* r);
    void unplace(Request* r);
    void release_all();
};

// ---------------------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------------------

int ghcm_server::init() {
    HIP_TRY(hipSetDevice(ordinal));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, ordinal));
    const int cus = p.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&up_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&down_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&ev_upload, hipEventDisableTiming));

    n_sel = cfg.n_active ? cfg.n_active
                         : (uint32_t)std::max(1.0, std::floor((double)cfg.compute_unit_overload * cus));
    // Waves far wider than the chip (the usual case for a serving loop) take the higher-
    // occupancy form of the tile kernel where one is built.
    const SolveKernel sk = get_solve_kernel((int)tw, (int)th, kind, cfg.kernel, false,
                                            (uint64_t)n_sel >= 4ull * (uint64_t)cus);
    launch = sk.launch;
    if (!launch)
        return fail(GHCM_ERR_UNSUPPORTED, "no kernel for tile %ux%u kind %d kernel %d", tw, th, kind,
                    cfg.kernel);
    const uint32_t resident = (uint32_t)std::max(1, sk.blocks_per_cu()) * (uint32_t)std::max(1, cus);
    grid = std::max(1u, std::min(n_sel, resident));
    lead = std::max<uint32_t>(2, cfg.waves_per_poll ? cfg.waves_per_poll : 4);
    step = cfg.launch_mode == 2 ? std::max<uint32_t>(1, lead / 2) : 1;

    const uint64_t T = max_tiles;
    const uint32_t shard_cap = (uint32_t)((T + kShards - 1) / kShards);
    int rc;
    if ((rc = arena.reserve(budget))) return rc;
    if ((rc = inst.reserve(sizeof(GhcmInst) * slots))) return rc;
    if ((rc = tinst.reserve(4 * T))) return rc;
    if ((rc = qwave.reserve(4 * T))) return rc;
    if ((rc = key.reserve(8 * T))) return rc;
    if ((rc = mask.reserve(8 * T))) return rc;
    if ((rc = tmin.reserve(4 * T))) return rc;
    if ((rc = list.reserve(8ull * kShards * shard_cap))) return rc;
    if ((rc = tstat.reserve(12 * T))) return rc;
    if ((rc = sel.reserve(4ull * std::max<uint64_t>(1, T)))) return rc;
    if ((rc = ctrl.reserve(sizeof(GhcmCtrl)))) return rc;
    if ((rc = sgen.reserve(4ull * slots))) return rc;
    if ((rc = totals.reserve(3 * sizeof(unsigned long long)))) return rc;
    if ((rc = rec.reserve(sizeof(HostRec)))) return rc;
    if (cfg.kernel == GHCM_KERNEL_SWEEP) {
        if ((rc = edge.reserve(8ull * th * T))) return rc;
        if ((rc = cedge.reserve(8ull * th * T))) return rc;
    }
    for (int k = 0; k < kSeedRing; ++k) {
        if ((rc = h_seeds[k].reserve(12 * 1024))) return rc;
        HIP_TRY(hipEventCreateWithFlags(&ev_seed[k], hipEventDisableTiming));
    }

    a.inst = inst.as<GhcmInst>();
    a.n_inst = slots;
    a.tinst = tinst.as<uint32_t>();
    a.qwave = qwave.as<uint32_t>();
    a.key = key.as<uint32_t>();
    a.mask = mask.as<uint32_t>();
    a.tmin = tmin.as<float>();
    a.list = list.as<uint32_t>();
    a.tiles = (uint32_t)T;
    a.shard_cap = shard_cap;
    a.tstat = tstat.as<uint32_t>();
    a.jobs = sel.as<uint32_t>();
    a.ctrl = ctrl.as<GhcmCtrl>();
    a.eps = cfg.min_change;
    a.iter_cap = cfg.inner_iter_cap;
    a.rec = rec.dev_as<HostRec>();
    a.sgen = sgen.as<uint32_t>();
    a.n_slots = slots;
    if (cfg.kernel == GHCM_KERNEL_SWEEP) {
        a.edge = edge.as<float>();
        a.cedge = cedge.as<float>();
        a.edge_h = th;
    }

    // Empty loop: lists empty, every slot generation 0 (no request carries it).
    std::memset(rec.p, 0, sizeof(HostRec));
    HIP_TRY(launch_solve_init(a, nullptr, 0, 0, n_sel, totals.as<unsigned long long>(), stream));
    HIP_TRY(launch_fill_u32(sgen.as<uint32_t>(), slots, 0u, stream));
    HIP_TRY(launch_fill_u32(a.qwave, T, 0u, stream));
    HIP_TRY(launch_fill_u32(const_cast<uint32_t*>(a.tinst), T, 0u, stream));
    HIP_TRY(launch_fill_u32(a.tstat, 3 * T, 0u, stream));
    HIP_TRY(hipStreamSynchronize(stream));

    if (cfg.launch_mode == 2) {
        // `step` waves as one HIP graph: the arguments never change for a server.
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (uint32_t k = 0; k < step && ce == hipSuccess; ++k) {
            ce = launch(a, grid, stream, nullptr, nullptr);
            if (ce == hipSuccess) ce = launch_wave_end(a, 0u, stream);
        }
        hipError_t ee = hipStreamEndCapture(stream, &graph);
        if (ce != hipSuccess || ee != hipSuccess)
            return fail(GHCM_ERR_HIP, "graph capture failed: %s",
                        hipGetErrorString(ce != hipSuccess ? ce : ee));
        hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            gexec = nullptr;
            return fail(GHCM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        }
    }

    arena_alloc.reset(budget);
    tile_alloc.reset(T);
    free_slots.clear();
    for (uint32_t s = slots; s-- > 0;) free_slots.push_back(s);
    live.assign(slots, nullptr);
    st.budget_bytes = budget;
    st.max_tiles = T;
    st.slots = slots;
    st.n_active = n_sel;
    st.waves_per_chunk = step;
    return GHCM_OK;
}

void ghcm_server::release_all() {
    for (hipStream_t s : {stream, up_stream, down_stream})
        if (s) (void)hipStreamSynchronize(s);
    for (DevBuf* b : {&arena, &inst, &tinst, &qwave, &key, &mask, &tmin, &list, &tstat, &sel, &ctrl,
                      &sgen, &totals, &edge, &cedge})
        b->release();
    for (int k = 0; k < kSeedRing; ++k) {
        h_seeds[k].release();
        if (ev_seed[k]) (void)hipEventDestroy(ev_seed[k]);
        ev_seed[k] = nullptr;
    }
    rec.release();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    ev_pool.clear();
    pool.release();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    if (ev_upload) (void)hipEventDestroy(ev_upload);
    ev_upload = nullptr;
    for (hipStream_t* s : {&stream, &up_stream, &down_stream}) {
        if (*s) (void)hipStreamDestroy(*s);
        *s = nullptr;
    }
}

// ---------------------------------------------------------------------------------------
// the server thread
// ---------------------------------------------------------------------------------------

int ghcm_server::get_event(hipEvent_t* ev) {
    if (ev_pool.empty()) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev_pool.push_back(e);
    }
    *ev = ev_pool.back();
    ev_pool.pop_back();
    return GHCM_OK;
}

// Placement of one request (caller holds mu): slot, tile range and arena bytes, or
// nothing at all.
bool ghcm_server::place(Request* r) {
    if (free_slots.empty()) return false;
    const uint64_t cells = (uint64_t)r->w * r->h;
    const uint64_t fb = r->device ? 0 : align_up(cells * kind_bytes(kind));
    const uint64_t tb = r->device ? 0 : align_up(cells * 4);
    uint64_t off = 0, tbase = 0;
    const uint64_t ntiles = (uint64_t)r->tiles_x * r->tiles_y;
    if (fb + tb && !arena_alloc.alloc(fb + tb, &off)) return false;
    if (!tile_alloc.alloc(ntiles, &tbase)) {
        if (fb + tb) arena_alloc.release(off, fb + tb);
        return false;
    }
    r->slot = free_slots.back();
    free_slots.pop_back();
    r->arena_off = off;
    r->arena_bytes = fb + tb;
    r->tile_base = tbase;
    r->tiles = ntiles;
    st.peak_bytes = std::max<uint64_t>(st.peak_bytes, arena_alloc.used());
    return true;
}

void ghcm_server::unplace(Request* r) {
    if (r->arena_bytes) arena_alloc.release(r->arena_off, r->arena_bytes);
    tile_alloc.release(r->tile_base, r->tiles);
    free_slots.push_back(r->slot);
    r->arena_bytes = r->tiles = 0;
}

// Admission of a batch of placed requests before the wave with tick `enqueued` + 1.
int ghcm_server::admit(std::vector<Request*>& batch, uint32_t enqueued) {
    char* base = arena.as<char>();
    size_t n_seeds = 0;
    bool uploads = false;
    for (Request* r : batch) {
        const uint64_t cells = (uint64_t)r->w * r->h;
        if (!r->device) {
            HIP_TRY(launch_copy(r->h_field_dev, base + r->arena_off, cells * kind_bytes(kind), up_stream));
            uploads = true;
        }
        n_seeds += r->seeds.size() / 2;
    }
    if (uploads) {
        HIP_TRY(hipEventRecord(ev_upload, up_stream));
        HIP_TRY(hipStreamWaitEvent(stream, ev_upload, 0));
    }
    // Seed staging: a ring of mapped buffers read by k_seed_serve; a buffer is reused once
    // the launch that read it, kSeedRing admissions ago, has run.
    const int ring = seed_next;
    if (n_seeds) {
        seed_next = (seed_next + 1) % kSeedRing;
        HIP_TRY(ghcm_capi::spin_wait(ev_seed[ring]));
        int rc = h_seeds[ring].reserve(12 * n_seeds);
        if (rc) return rc;
    }
    uint32_t* hs = h_seeds[ring].as<uint32_t>();
    size_t k = 0;
    for (Request* r : batch) {
        const uint64_t cells = (uint64_t)r->w * r->h;
        GhcmInst g{};
        if (r->device) {
            g.field = r->d_field;
            g.time = r->d_time;
        } else {
            g.field = base + r->arena_off;
            g.time = reinterpret_cast<float*>(base + r->arena_off + align_up(cells * kind_bytes(kind)));
        }
        g.w = r->w;
        g.h = r->h;
        g.tiles_x = r->tiles_x;
        g.tiles_y = r->tiles_y;
        g.tile_base = (uint32_t)r->tile_base;
        g.d = 1;
        g.tiles_z = 1;
        r->gen = next_gen++;
        if (next_gen == 0) next_gen = 1;
        // The slot's entry, tile states and tile map, its time field at +INF and its tiles'
        // edge columns (sweep kernel) in one launch.
        HIP_TRY(launch_admit_tiles(a, g, r->slot, r->gen, stream, tw, th, kind));
        for (size_t s = 0; s + 1 < r->seeds.size(); s += 2) {
            hs[k++] = r->seeds[s];
            hs[k++] = r->seeds[s + 1];
            hs[k++] = r->slot;
        }
        r->admit_tick = enqueued + 1;
        r->state = ReqState::Live;
        live[r->slot] = r;
    }
    if (n_seeds) {
        HIP_TRY(launch_seed_serve(a, h_seeds[ring].dev_as<uint32_t>(), (uint32_t)n_seeds, tw, th, stream));
        HIP_TRY(hipEventRecord(ev_seed[ring], stream));
    }
    HIP_TRY(launch_wake(a.ctrl, stream));
    return GHCM_OK;
}

int ghcm_server::enqueue_step() {
    if (gexec) {
        HIP_TRY(hipGraphLaunch(gexec, stream));
    } else {
        HIP_TRY(launch(a, grid, stream, nullptr, nullptr));
        HIP_TRY(launch_wave_end(a, 0u, stream));
    }
    return GHCM_OK;
}

int ghcm_server::loop() {
    HIP_TRY(hipSetDevice(ordinal));
    volatile HostRec* hr = rec.as<HostRec>();
    std::vector<Request*> finishing;
    uint32_t enqueued = 0;               // wave boundaries launched (modulo 2^32)
    uint32_t scanned_tick = 0;
    uint32_t n_live = 0;
    uint64_t idle_spins = 0;
    char* base = arena.as<char>();
    for (;;) {
        bool progressed = false;
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop && !drain) return GHCM_OK;
                if (n_live > 0 || !finishing.empty() || !pending.empty()) break;
                if (stop) return GHCM_OK;           // drained
                cv_work.wait(lk);
            }
            while (!pending.empty() && place(pending.front())) {
                batch.push_back(pending.front());
                pending.pop_front();
            }
            st.admitted += batch.size();
        }
        if (!batch.empty()) {
            const auto now = Clock::now();
            for (Request* r : batch) r->t_admit = now;
            int rc = admit(batch, enqueued);
            if (rc) return rc;
            n_live += (uint32_t)batch.size();
            progressed = true;
            std::lock_guard<std::mutex> lk(mu);
            st.peak_live = std::max<uint64_t>(st.peak_live, n_live);
        }
        const uint32_t ticks = (uint32_t)hr->progress;
        // Keep `lead` waves queued ahead of the last executed boundary while work is live.
        if (n_live > 0 && (uint32_t)(enqueued - ticks) < lead) {
            int rc = enqueue_step();
            if (rc) return rc;
            enqueued += step;
            progressed = true;
            std::lock_guard<std::mutex> lk(mu);
            st.waves_enqueued += step;
        }
        // Completions: each boundary reports (generation, queued tiles) per slot.
        if (n_live > 0 && ticks != scanned_tick) {
            scanned_tick = ticks;
            std::vector<Request*> done_now;
            for (uint32_t s = 0; s < slots; ++s) {
                Request* r = live[s];
                if (!r) continue;
                const unsigned long long v = hr->slot[s];
                if ((uint32_t)(v >> 32) != r->gen || (uint32_t)v != 0u) continue;
                live[s] = nullptr;
                --n_live;
                r->waves = (uint32_t)(ticks - r->admit_tick) + 1u;
                r->state = ReqState::Copying;
                if (!r->device) {
                    // Times are final; the copy runs behind the waves already queued.
                    int rc = get_event(&r->ev_ready);
                    if (!rc) rc = get_event(&r->ev_done);
                    if (rc) return rc;
                    const uint64_t cells = (uint64_t)r->w * r->h;
                    const char* tsrc = base + r->arena_off + align_up(cells * kind_bytes(kind));
                    HIP_TRY(hipEventRecord(r->ev_ready, stream));
                    HIP_TRY(hipStreamWaitEvent(down_stream, r->ev_ready, 0));
                    HIP_TRY(launch_copy(tsrc, r->h_time_dev, cells * 4, down_stream));
                    HIP_TRY(hipEventRecord(r->ev_done, down_stream));
                }
                finishing.push_back(r);
                progressed = true;
            }
            std::lock_guard<std::mutex> lk(mu);
            st.busy_waves = hr->busy_waves;
            st.loop_wave = hr->wave;
        }
        // Requests whose results are in place are handed back; their resources freed.
        if (!finishing.empty()) {
            std::vector<Request*> keep, done;
            for (Request* r : finishing) {
                if (r->ev_done) {
                    const hipError_t q = hipEventQuery(r->ev_done);
                    if (q == hipErrorNotReady) { keep.push_back(r); continue; }
                    if (q != hipSuccess) return fail(GHCM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
                    ev_pool.push_back(r->ev_ready);
                    ev_pool.push_back(r->ev_done);
                    r->ev_ready = r->ev_done = nullptr;
                }
                done.push_back(r);
            }
            finishing.swap(keep);
            if (!done.empty()) {
                progressed = true;
                const auto now = Clock::now();
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : done) {
                    unplace(r);
                    if (r->h_field) {
                        pool.put(r->h_field, (size_t)r->w * r->h * kind_bytes(kind));
                        r->h_field = nullptr;
                    }
                    r->t_done = now;
                    r->status = GHCM_OK;
                    r->state = ReqState::Done;
                    ++st.completed;
                }
                cv_done.notify_all();
            }
        }
        if (progressed) {
            idle_spins = 0;
        } else if (++idle_spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(stream);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(GHCM_ERR_HIP, "serving loop: %s", hipGetErrorString(q));
            std::this_thread::yield();
        } else {
            __builtin_ia32_pause();
        }
    }
}

void ghcm_server::run() {
    const int rc = loop();
    // No copy may still be writing into a caller's buffer once its request is reported.
    for (hipStream_t s : {down_stream, up_stream, stream})
        if (s) (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(mu);
    if (rc != GHCM_OK) {
        fatal = rc;
        fatal_msg = ghcm_last_error();
    }
    // Whatever did not finish ends with the loop's status (or STOPPED).
    const int end_status = rc != GHCM_OK ? rc : GHCM_ERR_STOPPED;
    for (auto& kv : tickets) {
        Request* r = kv.second;
        if (r->state != ReqState::Done) {
            r->state = ReqState::Done;
            r->status = end_status;
            r->t_done = Clock::now();
        }
    }
    pending.clear();
    exited = true;
    cv_done.notify_all();
}

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------

extern "C" void ghcm_serve_config_default(ghcm_serve_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->hbm_budget_bytes = 8ull << 30;
    c->max_slots = 256;
    c->field_kind = GHCM_FIELD_COST_F32;
}

extern "C" int ghcm_serve_start(int ordinal, const ghcm_config* cfg, const ghcm_serve_config* scfg,
                                ghcm_server** out) {
    if (!out) return fail(GHCM_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    ghcm_config c;
    if (cfg) c = *cfg; else ghcm_config_default(&c);
    ghcm_serve_config sc;
    if (scfg) sc = *scfg; else ghcm_serve_config_default(&sc);
    int rc = ghcm_capi::check_config(&c);
    if (rc) return rc;
    if (c.scheduler != GHCM_SCHED_DEVICE_TOPN)
        return fail(GHCM_ERR_UNSUPPORTED, "serving runs the device top-n scheduler only");
    if (!ghcm_capi::valid_kind(sc.field_kind))
        return fail(GHCM_ERR_INVALID_ARG, "unknown field kind %d", sc.field_kind);
    const uint32_t slots = sc.max_slots ? sc.max_slots : 256;
    if (slots > kMaxServeSlots)
        return fail(GHCM_ERR_INVALID_ARG, "max_slots %u above %u", slots, kMaxServeSlots);
    const uint64_t budget = sc.hbm_budget_bytes ? sc.hbm_budget_bytes : (8ull << 30);
    const uint64_t per_tile = (uint64_t)c.tile_w * c.tile_h * (kind_bytes(sc.field_kind) + 4);
    uint64_t max_tiles = sc.max_tiles ? sc.max_tiles : budget / per_tile + 64ull * slots;
    max_tiles = std::min<uint64_t>(max_tiles, 1ull << 28);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ordinal < 0 || ordinal >= n)
        return fail(GHCM_ERR_NO_DEVICE, "device %d not present (%d devices)", ordinal, n);
    ghcm_server* s = new (std::nothrow) ghcm_server();
    if (!s) return fail(GHCM_ERR_OUT_OF_MEMORY, "server allocation failed");
    s->ordinal = ordinal;
    s->cfg = c;
    s->scfg = sc;
    s->kind = sc.field_kind;
    s->tw = c.tile_w;
    s->th = c.tile_h;
    s->slots = slots;
    s->budget = budget;
    s->max_tiles = (uint32_t)max_tiles;
    rc = s->init();
    if (rc) {
        const std::string msg = ghcm_last_error();
        s->release_all();
        delete s;
        return fail(rc, "%s", msg.c_str());
    }
    try {
        s->thread = std::thread([s] { s->run(); });
    } catch (...) {
        s->release_all();
        delete s;
        return fail(GHCM_ERR_OUT_OF_MEMORY, "could not start the server thread");
    }
    *out = s;
    return GHCM_OK;
}

namespace {

// Checks shared by both submit forms; fills the request's size and tile counts.
int check_request(ghcm_server* s, uint32_t w, uint32_t h, const uint32_t* seeds_xy,
                  uint32_t n_seeds, bool device, Request* r) {
    if (w == 0 || h == 0) return fail(GHCM_ERR_INVALID_ARG, "zero-sized field");
    if (w > GHCM_MAX_WIDTH) return fail(GHCM_ERR_INVALID_ARG, "width %u above %u", w, GHCM_MAX_WIDTH);
    if (n_seeds && !seeds_xy) return fail(GHCM_ERR_INVALID_ARG, "seeds are null");
    for (uint32_t k = 0; k < n_seeds; ++k)
        if (seeds_xy[2 * k] >= w || seeds_xy[2 * k + 1] >= h)
            return fail(GHCM_ERR_INVALID_ARG, "seed %u (%u,%u) outside %ux%u", k, seeds_xy[2 * k],
                        seeds_xy[2 * k + 1], w, h);
    const uint64_t cells = (uint64_t)w * h;
    const uint32_t tx = (w + s->tw - 1) / s->tw, ty = (h + s->th - 1) / s->th;
    const uint64_t bytes = device ? 0 : align_up(cells * kind_bytes(s->kind)) + align_up(cells * 4);
    if (bytes > s->budget)
        return fail(GHCM_ERR_OVER_BUDGET, "request needs %llu bytes, budget %llu",
                    (unsigned long long)bytes, (unsigned long long)s->budget);
    if ((uint64_t)tx * ty > s->max_tiles)
        return fail(GHCM_ERR_OVER_BUDGET, "request needs %llu tiles, capacity %u",
                    (unsigned long long)tx * ty, s->max_tiles);
    r->device = device;
    r->w = w;
    r->h = h;
    r->tiles_x = tx;
    r->tiles_y = ty;
    try {
        r->seeds.assign(seeds_xy, seeds_xy + 2ull * n_seeds);
    } catch (const std::bad_alloc&) {
        return fail(GHCM_ERR_OUT_OF_MEMORY, "%u seeds", n_seeds);
    }
    return GHCM_OK;
}

int enqueue_request(ghcm_server* s, Request* r, uint64_t* ticket) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop || s->exited) {
            s->pool.put(r->h_field, (size_t)r->w * r->h * kind_bytes(s->kind));
            s->pool.put(r->h_time, (size_t)r->w * r->h * 4);
            delete r;
            return s->fatal ? fail(s->fatal, "server failed: %s", s->fatal_msg.c_str())
                            : fail(GHCM_ERR_STOPPED, "server is stopping");
        }
        r->ticket = s->next_ticket;
        try {
            s->tickets[r->ticket] = r;
            s->pending.push_back(r);
        } catch (const std::bad_alloc&) {
            s->tickets.erase(r->ticket);
            s->pool.put(r->h_field, (size_t)r->w * r->h * kind_bytes(s->kind));
            s->pool.put(r->h_time, (size_t)r->w * r->h * 4);
            delete r;
            return fail(GHCM_ERR_OUT_OF_MEMORY, "request queue allocation failed");
        }
        ++s->next_ticket;
        ++s->st.submitted;
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return GHCM_OK;
}

}  // namespace

extern "C" int ghcm_serve_submit(ghcm_server* s, const void* field, uint32_t w, uint32_t h,
                                 const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
                                 uint64_t* ticket) {
  // The staging copy allocates (the copy pool's queue): nothing may escape the C boundary.
  return no_throw([&]() -> int {
    if (!s) return fail(GHCM_ERR_INVALID_ARG, "server is null");
    if (!field || !time_out || !ticket) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    // The pinned staging below is mapped for, and addressed on, the server's device.
    HIP_TRY(hipSetDevice(s->ordinal));
    Request* r = new (std::nothrow) Request();
    if (!r) return fail(GHCM_ERR_OUT_OF_MEMORY, "request allocation failed");
    int rc = check_request(s, w, h, seeds_xy, n_seeds, false, r);
    if (rc) {
        delete r;
        return rc;
    }
    r->t_submit = Clock::now();
    // The field is staged in pinned memory here, on the caller's thread, so the server
    // thread only issues asynchronous copies.
    const size_t fbytes = (size_t)w * h * kind_bytes(s->kind), tbytes = (size_t)w * h * 4;
    r->h_field = s->pool.get(fbytes, &r->h_field_dev);
    r->h_time = s->pool.get(tbytes, &r->h_time_dev);
    auto drop = [&] {
        s->pool.put(r->h_field, fbytes);
        s->pool.put(r->h_time, tbytes);
        delete r;
    };
    if (!r->h_field || !r->h_time) {
        drop();
        return fail(GHCM_ERR_OUT_OF_MEMORY, "pinned staging of %zu bytes failed", fbytes + tbytes);
    }
    try {
        par_memcpy(r->h_field, field, fbytes);
    } catch (...) {
        drop();
        throw;
    }
    r->time_out = time_out;
    return enqueue_request(s, r, ticket);
  });
}

extern "C" int ghcm_serve_submit_device(ghcm_server* s, const void* d_field, uint32_t w,
                                        uint32_t h, const uint32_t* seeds_xy, uint32_t n_seeds,
                                        float* d_time, uint64_t* ticket) {
    if (!s) return fail(GHCM_ERR_INVALID_ARG, "server is null");
    if (!d_field || !d_time || !ticket) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    Request* r = new (std::nothrow) Request();
    if (!r) return fail(GHCM_ERR_OUT_OF_MEMORY, "request allocation failed");
    int rc = check_request(s, w, h, seeds_xy, n_seeds, true, r);
    if (rc) {
        delete r;
        return rc;
    }
    r->t_submit = Clock::now();
    r->d_field = d_field;
    r->d_time = d_time;
    return enqueue_request(s, r, ticket);
}

extern "C" int ghcm_serve_wait(ghcm_server* s, uint64_t ticket, double timeout_s,
                               ghcm_serve_result* out) {
  return no_throw([&]() -> int {
    if (!s) return fail(GHCM_ERR_INVALID_ARG, "server is null");
    std::unique_lock<std::mutex> lk(s->mu);
    // ghcm_serve_stop frees the server only once no wait is inside this function.
    struct Leave {
        ghcm_server* s;
        std::unique_lock<std::mutex>& lk;
        ~Leave() {
            if (!lk.owns_lock()) lk.lock();
            if (--s->waiters == 0) s->cv_done.notify_all();
        }
    };
    ++s->waiters;
    Leave leave{s, lk};
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end())
        return fail(GHCM_ERR_NOT_FOUND, "ticket %llu unknown or already collected", (unsigned long long)ticket);
    Request* r = it->second;
    if (r->waited)
        return fail(GHCM_ERR_INVALID_ARG, "ticket %llu is already being waited on", (unsigned long long)ticket);
    r->waited = true;
    auto ready = [r] { return r->state == ReqState::Done; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        r->waited = false;
        return fail(GHCM_ERR_TIMEOUT, "ticket %llu still running", (unsigned long long)ticket);
    }
    s->tickets.erase(ticket);
    const std::string fatal_msg = s->fatal_msg;
    lk.unlock();
    int status = r->status;
    // Host request: the result lands in the caller's buffer from the pinned staging.  The
    // request is collected (its ticket erased) whatever happens here, so a failed copy is
    // reported and the request still freed below.
    if (!r->device && status == GHCM_OK) {
        const int rc_copy = no_throw([&] {
            par_memcpy(r->time_out, r->h_time, (size_t)r->w * r->h * 4);
            return (int)GHCM_OK;
        });
        if (rc_copy != GHCM_OK) status = rc_copy;
    }
    const auto t_out = Clock::now();
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->status = status;
        out->w = r->w;
        out->h = r->h;
        out->waves = r->waves;
        const bool admitted = r->t_admit.time_since_epoch().count() != 0;
        out->queue_us = admitted ? std::chrono::duration<double, std::micro>(r->t_admit - r->t_submit).count() : 0.0;
        out->latency_us = std::chrono::duration<double, std::micro>(
            (r->device || status != GHCM_OK ? r->t_done : t_out) - r->t_submit).count();
    }
    s->pool.put(r->h_field, (size_t)r->w * r->h * kind_bytes(s->kind));
    s->pool.put(r->h_time, (size_t)r->w * r->h * 4);
    delete r;
    if (status != GHCM_OK)
        return fail(status, "request %llu: %s", (unsigned long long)ticket,
                    fatal_msg.empty() ? ghcm_status_string(status) : fatal_msg.c_str());
    return GHCM_OK;
  });
}

extern "C" int ghcm_serve_get_stats(ghcm_server* s, ghcm_serve_stats* out) {
    if (!s || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->pending = s->pending.size();
    out->bytes_in_use = s->arena_alloc.used();
    out->waiters = s->waiters;
    return GHCM_OK;
}

extern "C" int ghcm_serve_stop(ghcm_server* s, int drain) {
    if (!s) return GHCM_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->drain = drain != 0;
    }
    s->cv_work.notify_all();
    if (s->thread.joinable()) s->thread.join();
    {
        // Every request is Done now, so each wait still inside returns promptly.
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [s] { return s->waiters == 0; });
    }
    (void)hipSetDevice(s->ordinal);
    s->release_all();
    const int fatal = s->fatal;
    const std::string msg = s->fatal_msg;
    for (auto& kv : s->tickets) delete kv.second;
    delete s;
    if (fatal) return fail(fatal, "server failed: %s", msg.c_str());
    return GHCM_OK;
}
