// serve.cpp -- continuous-batching solve server (ghcm_serve_*).
//
// One device wave loop runs for the lifetime of the server and never restarts between
// requests.  A request is admitted at a wave boundary when an instance slot, a range of
// tile state and room for its field and time buffer inside the HBM arena are all free
// (first come, first served).  Admission is three launches on the loop's stream: the
// slot's instance entry and tile range are reset (k_admit_tiles), the time buffer is set
// to +INF, and the seeds queue their tiles for the next wave (k_seed_serve), exactly as
// a plain solve's seeds are queued for its first wave (src/eikonal_hcm.cpp:189-238).
// From then on the request's tiles compete in the same top-n selection as every other
// live request's.  Each wave boundary counts the queued tiles per slot (k_wave_end);
// the host reads those counts once per chunk of waves, and a request whose count is 0
// at a boundary after its admission is finished: tiles are only ever queued by seeds or
// by a solved neighbour of the same instance, so nothing can reopen it.  Its times are
// copied out on a second stream, and its slot, tiles and arena bytes are freed once the
// copy has landed.
//
// The reference solves one field per call (src/eikonal_hcm.cpp:186-529); this loop is
// the serving form of the same wave algorithm.
#include <hip/hip_runtime.h>

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
using ghcm_capi::fail;
using ghcm_capi::kind_bytes;

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

constexpr uint64_t kAlign = 256;
inline uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

enum class ReqState { Pending, Live, Copying, Done };

struct Request {
    uint64_t ticket = 0;
    const void* field = nullptr;
    uint32_t w = 0, h = 0;
    std::vector<uint32_t> seeds;        // (x, y) pairs
    float* time_out = nullptr;
    ReqState state = ReqState::Pending;
    int status = GHCM_OK;
    Clock::time_point t_submit, t_admit, t_done;
    // placement
    uint32_t slot = 0;
    uint64_t arena_off = 0, arena_bytes = 0;
    uint64_t tile_base = 0, tiles = 0;
    uint32_t tiles_x = 0, tiles_y = 0;
    uint64_t admit_seq = 0;             // first chunk enqueued after admission
    uint64_t waves = 0;
    hipEvent_t ev_done = nullptr;       // result copy landed
};

struct SnapHead {                        // first words of GhcmCtrl read back per chunk
    uint32_t wave, cur, done, n_select, n_jobs, jobs_from_list, busy_waves, total;
};

}  // namespace

struct ghcm_server {
    static constexpr int kSlots = 4;     // chunks in flight

    int ordinal = 0;
    ghcm_config cfg{};
    ghcm_serve_config scfg{};
    int kind = GHCM_FIELD_COST_F32;
    uint32_t tw = 0, th = 0, n_sel = 0, grid = 1, chunk = 4, slots = 0, max_tiles = 0;
    uint64_t budget = 0;
    solve_launcher_t launch = nullptr;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    hipEvent_t ev_upload = nullptr;
    hipEvent_t ev_slot[kSlots] = {};
    uint64_t slot_seq[kSlots] = {};
    hipGraphExec_t gexec = nullptr;
    DevBuf arena, inst, tinst, qwave, key, mask, tmin, list, tstat, sel, ctrl, iact;
    DevBuf d_seeds[kSlots];
    PinnedBuf h_seeds[kSlots], h_snap[kSlots];
    SolveArgs a{};
    RangeAlloc arena_alloc, tile_alloc;
    std::vector<uint32_t> free_slots;
    std::vector<Request*> live;          // by slot (server thread only)
    std::vector<hipEvent_t> ev_pool;     // server thread only

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> tickets;
    uint64_t next_ticket = 1;
    bool stop = false, drain = true, exited = false;
    int fatal = GHCM_OK;
    std::string fatal_msg;
    ghcm_serve_stats st{};
    std::thread thread;

    int init();
    void run();
    int loop();
    int admit(std::vector<Request*>& batch, int head, uint64_t seq);
    int enqueue_chunk(int head);
    int harvest(int sl, std::vector<Request*>& finishing);
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
    HIP_TRY(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&ev_upload, hipEventDisableTiming));
    for (int k = 0; k < kSlots; ++k) HIP_TRY(hipEventCreateWithFlags(&ev_slot[k], hipEventDisableTiming));

    const SolveKernel sk = get_solve_kernel((int)tw, (int)th, kind, cfg.kernel, false);
    launch = sk.launch;
    if (!launch)
        return fail(GHCM_ERR_UNSUPPORTED, "no kernel for tile %ux%u kind %d kernel %d", tw, th, kind,
                    cfg.kernel);
    n_sel = cfg.n_active ? cfg.n_active
                         : (uint32_t)std::max(1.0, std::floor((double)cfg.compute_unit_overload * cus));
    const uint32_t resident = (uint32_t)std::max(1, sk.blocks_per_cu()) * (uint32_t)std::max(1, cus);
    grid = std::max(1u, std::min(n_sel, resident));
    chunk = cfg.waves_per_poll ? cfg.waves_per_poll : 4;

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
    if ((rc = sel.reserve(4ull * std::max<uint64_t>(1, std::min<uint64_t>(n_sel, T))))) return rc;
    if ((rc = ctrl.reserve(sizeof(GhcmCtrl)))) return rc;
    if ((rc = iact.reserve(4ull * slots))) return rc;
    for (int k = 0; k < kSlots; ++k)
        if ((rc = h_snap[k].reserve(sizeof(SnapHead) + 4ull * slots))) return rc;

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
    a.iact = iact.as<uint32_t>();
    a.n_iact = slots;

    // Idle loop: empty lists, `done` set, every slot's count 0.
    if ((rc = h_snap[0].reserve(std::max<size_t>(sizeof(GhcmCtrl), sizeof(SnapHead) + 4ull * slots))))
        return rc;
    GhcmCtrl* h0 = h_snap[0].as<GhcmCtrl>();
    std::memset(h0, 0, sizeof(GhcmCtrl));
    h0->done = 1;
    h0->n_select = n_sel;
    HIP_TRY(hipMemcpyAsync(a.ctrl, h0, sizeof(GhcmCtrl), hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_fill_u32(a.iact, slots, 0u, stream));
    HIP_TRY(launch_fill_u32(a.qwave, T, 0u, stream));
    HIP_TRY(launch_fill_u32(const_cast<uint32_t*>(a.tinst), T, 0u, stream));
    HIP_TRY(launch_fill_u32(a.tstat, 3 * T, 0u, stream));
    HIP_TRY(hipStreamSynchronize(stream));

    if (cfg.launch_mode == 2) {
        // One chunk of waves as a HIP graph: the arguments never change for a server.
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (uint32_t k = 0; k < chunk && ce == hipSuccess; ++k) {
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
    st.waves_per_chunk = chunk;
    return GHCM_OK;
}

void ghcm_server::release_all() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (DevBuf* b : {&arena, &inst, &tinst, &qwave, &key, &mask, &tmin, &list, &tstat, &sel, &ctrl, &iact})
        b->release();
    for (int k = 0; k < kSlots; ++k) {
        d_seeds[k].release();
        h_seeds[k].release();
        h_snap[k].release();
        if (ev_slot[k]) (void)hipEventDestroy(ev_slot[k]);
        ev_slot[k] = nullptr;
    }
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    ev_pool.clear();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    if (ev_upload) (void)hipEventDestroy(ev_upload);
    ev_upload = nullptr;
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (stream) (void)hipStreamDestroy(stream);
    copy_stream = stream = nullptr;
}

// ---------------------------------------------------------------------------------------
// the server thread
// ---------------------------------------------------------------------------------------

// Placement of one request (caller holds mu): slot, tile range and arena bytes, or
// nothing at all.
bool ghcm_server::place(Request* r) {
    if (free_slots.empty()) return false;
    const uint64_t cells = (uint64_t)r->w * r->h;
    const uint64_t fb = align_up(cells * kind_bytes(kind)), tb = align_up(cells * 4);
    uint64_t off = 0, tbase = 0;
    const uint64_t ntiles = (uint64_t)r->tiles_x * r->tiles_y;
    if (!arena_alloc.alloc(fb + tb, &off)) return false;
    if (!tile_alloc.alloc(ntiles, &tbase)) {
        arena_alloc.release(off, fb + tb);
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
    arena_alloc.release(r->arena_off, r->arena_bytes);
    tile_alloc.release(r->tile_base, r->tiles);
    free_slots.push_back(r->slot);
    r->arena_bytes = r->tiles = 0;
}

// Admission of a batch of placed requests, enqueued before chunk `seq` (slot `head`).
int ghcm_server::admit(std::vector<Request*>& batch, int head, uint64_t seq) {
    char* base = arena.as<char>();
    size_t n_seeds = 0;
    for (Request* r : batch) {
        const uint64_t cells = (uint64_t)r->w * r->h;
        HIP_TRY(hipMemcpyAsync(base + r->arena_off, r->field, cells * kind_bytes(kind),
                               hipMemcpyHostToDevice, copy_stream));
        n_seeds += r->seeds.size() / 2;
    }
    HIP_TRY(hipEventRecord(ev_upload, copy_stream));
    HIP_TRY(hipStreamWaitEvent(stream, ev_upload, 0));
    int rc;
    if ((rc = h_seeds[head].reserve(12 * std::max<size_t>(1, n_seeds)))) return rc;
    if ((rc = d_seeds[head].reserve(12 * std::max<size_t>(1, n_seeds)))) return rc;
    uint32_t* hs = h_seeds[head].as<uint32_t>();
    size_t k = 0;
    for (Request* r : batch) {
        const uint64_t cells = (uint64_t)r->w * r->h;
        GhcmInst g{};
        g.field = base + r->arena_off;
        g.time = reinterpret_cast<float*>(base + r->arena_off + align_up(cells * kind_bytes(kind)));
        g.w = r->w;
        g.h = r->h;
        g.tiles_x = r->tiles_x;
        g.tiles_y = r->tiles_y;
        g.tile_base = (uint32_t)r->tile_base;
        HIP_TRY(launch_admit_tiles(a, g, r->slot, stream));
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(g.time), cells, kInfBits, stream));
        for (size_t s = 0; s + 1 < r->seeds.size(); s += 2) {
            hs[k++] = r->seeds[s];
            hs[k++] = r->seeds[s + 1];
            hs[k++] = r->slot;
        }
        r->admit_seq = seq;
        r->state = ReqState::Live;
        live[r->slot] = r;
    }
    if (n_seeds) {
        HIP_TRY(hipMemcpyAsync(d_seeds[head].p, hs, 12 * n_seeds, hipMemcpyHostToDevice, stream));
        HIP_TRY(launch_seed_serve(a, d_seeds[head].as<uint32_t>(), (uint32_t)n_seeds, tw, th, stream));
    }
    return GHCM_OK;
}

int ghcm_server::enqueue_chunk(int head) {
    if (gexec) {
        HIP_TRY(hipGraphLaunch(gexec, stream));
    } else {
        for (uint32_t k = 0; k < chunk; ++k) {
            HIP_TRY(launch(a, grid, stream, nullptr, nullptr));
            HIP_TRY(launch_wave_end(a, 0u, stream));
        }
    }
    char* snap = h_snap[head].as<char>();
    HIP_TRY(hipMemcpyAsync(snap, a.ctrl, sizeof(SnapHead), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(snap + sizeof(SnapHead), a.iact, 4ull * slots, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(ev_slot[head], stream));
    return GHCM_OK;
}

// Reads the snapshot of a completed chunk: requests with no queued tile are finished.
int ghcm_server::harvest(int sl, std::vector<Request*>& finishing) {
    const char* snap = h_snap[sl].as<char>();
    SnapHead hd;
    std::memcpy(&hd, snap, sizeof hd);
    const uint32_t* cnt = reinterpret_cast<const uint32_t*>(snap + sizeof(SnapHead));
    const uint64_t seq = slot_seq[sl];
    {
        std::lock_guard<std::mutex> lk(mu);
        st.busy_waves = hd.busy_waves;
        st.loop_wave = hd.wave;
    }
    char* base = arena.as<char>();
    for (uint32_t s = 0; s < slots; ++s) {
        Request* r = live[s];
        if (!r || r->admit_seq > seq || cnt[s] != 0u) continue;
        live[s] = nullptr;
        r->waves = (seq - r->admit_seq + 1) * chunk;
        if (ev_pool.empty()) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev_pool.push_back(e);
        }
        r->ev_done = ev_pool.back();
        ev_pool.pop_back();
        const uint64_t cells = (uint64_t)r->w * r->h;
        const char* tsrc = base + r->arena_off + align_up(cells * kind_bytes(kind));
        HIP_TRY(hipMemcpyAsync(r->time_out, tsrc, cells * 4, hipMemcpyDeviceToHost, copy_stream));
        HIP_TRY(hipEventRecord(r->ev_done, copy_stream));
        r->state = ReqState::Copying;
        finishing.push_back(r);
    }
    return GHCM_OK;
}

int ghcm_server::loop() {
    HIP_TRY(hipSetDevice(ordinal));
    std::vector<Request*> finishing;
    uint64_t seq = 0;
    int inflight = 0, head = 0;
    uint32_t n_live = 0;
    for (;;) {
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop && !drain) return GHCM_OK;
                const bool busy = n_live > 0 || !finishing.empty() || inflight > 0;
                if (busy || !pending.empty()) break;
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
            int rc = admit(batch, head, seq);
            if (rc) return rc;
            n_live += (uint32_t)batch.size();
            std::lock_guard<std::mutex> lk(mu);
            st.peak_live = std::max<uint64_t>(st.peak_live, n_live);
        }
        if (n_live > 0) {
            int rc = enqueue_chunk(head);
            if (rc) return rc;
            slot_seq[head] = seq++;
            ++inflight;
            head = (head + 1) % kSlots;
            std::lock_guard<std::mutex> lk(mu);
            st.waves_enqueued += chunk;
        }
        // Harvest completed chunks: always the oldest when the ring is full or nothing
        // more will be launched, otherwise only those already finished.
        while (inflight > 0) {
            const int oldest = (head - inflight + kSlots) % kSlots;
            if (inflight == kSlots || n_live == 0) {
                HIP_TRY(hipEventSynchronize(ev_slot[oldest]));
            } else {
                const hipError_t q = hipEventQuery(ev_slot[oldest]);
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess) return fail(GHCM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
            }
            const size_t before = finishing.size();
            int rc = harvest(oldest, finishing);
            if (rc) return rc;
            n_live -= (uint32_t)(finishing.size() - before);
            --inflight;
        }
        // Results whose copy has landed are handed back and their resources freed.
        if (!finishing.empty()) {
            if (n_live == 0 && inflight == 0) HIP_TRY(hipEventSynchronize(finishing.front()->ev_done));
            std::vector<Request*> keep;
            std::vector<Request*> done;
            for (Request* r : finishing) {
                const hipError_t q = hipEventQuery(r->ev_done);
                if (q == hipErrorNotReady) { keep.push_back(r); continue; }
                if (q != hipSuccess) return fail(GHCM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
                ev_pool.push_back(r->ev_done);
                r->ev_done = nullptr;
                done.push_back(r);
            }
            finishing.swap(keep);
            if (!done.empty()) {
                const auto now = Clock::now();
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : done) {
                    unplace(r);
                    r->t_done = now;
                    r->status = GHCM_OK;
                    r->state = ReqState::Done;
                    ++st.completed;
                }
                cv_done.notify_all();
            }
        }
    }
}

void ghcm_server::run() {
    const int rc = loop();
    // No copy may still be writing into a caller's buffer once its request is reported.
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (stream) (void)hipStreamSynchronize(stream);
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

extern "C" int ghcm_serve_submit(ghcm_server* s, const void* field, uint32_t w, uint32_t h,
                                 const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
                                 uint64_t* ticket) {
    if (!s) return fail(GHCM_ERR_INVALID_ARG, "server is null");
    if (!field || !time_out || !ticket) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    if (w == 0 || h == 0) return fail(GHCM_ERR_INVALID_ARG, "zero-sized field");
    if (n_seeds && !seeds_xy) return fail(GHCM_ERR_INVALID_ARG, "seeds are null");
    for (uint32_t k = 0; k < n_seeds; ++k)
        if (seeds_xy[2 * k] >= w || seeds_xy[2 * k + 1] >= h)
            return fail(GHCM_ERR_INVALID_ARG, "seed %u (%u,%u) outside %ux%u", k, seeds_xy[2 * k],
                        seeds_xy[2 * k + 1], w, h);
    const uint64_t cells = (uint64_t)w * h;
    const uint32_t tx = (w + s->tw - 1) / s->tw, ty = (h + s->th - 1) / s->th;
    const uint64_t bytes = align_up(cells * kind_bytes(s->kind)) + align_up(cells * 4);
    if (bytes > s->budget)
        return fail(GHCM_ERR_OVER_BUDGET, "request needs %llu bytes, budget %llu",
                    (unsigned long long)bytes, (unsigned long long)s->budget);
    if ((uint64_t)tx * ty > s->max_tiles)
        return fail(GHCM_ERR_OVER_BUDGET, "request needs %llu tiles, capacity %u",
                    (unsigned long long)tx * ty, s->max_tiles);
    Request* r = new (std::nothrow) Request();
    if (!r) return fail(GHCM_ERR_OUT_OF_MEMORY, "request allocation failed");
    r->field = field;
    r->w = w;
    r->h = h;
    r->tiles_x = tx;
    r->tiles_y = ty;
    r->time_out = time_out;
    r->seeds.assign(seeds_xy, seeds_xy + 2ull * n_seeds);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop || s->exited) {
            delete r;
            return s->fatal ? fail(s->fatal, "server failed: %s", s->fatal_msg.c_str())
                            : fail(GHCM_ERR_STOPPED, "server is stopping");
        }
        r->ticket = s->next_ticket++;
        s->tickets[r->ticket] = r;
        s->pending.push_back(r);
        ++s->st.submitted;
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return GHCM_OK;
}

extern "C" int ghcm_serve_wait(ghcm_server* s, uint64_t ticket, double timeout_s,
                               ghcm_serve_result* out) {
    if (!s) return fail(GHCM_ERR_INVALID_ARG, "server is null");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end())
        return fail(GHCM_ERR_NOT_FOUND, "ticket %llu unknown or already collected", (unsigned long long)ticket);
    Request* r = it->second;
    auto ready = [r] { return r->state == ReqState::Done; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return fail(GHCM_ERR_TIMEOUT, "ticket %llu still running", (unsigned long long)ticket);
    }
    const int status = r->status;
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->status = status;
        out->w = r->w;
        out->h = r->h;
        out->waves = r->waves;
        const bool admitted = r->t_admit.time_since_epoch().count() != 0;
        out->queue_us = admitted ? std::chrono::duration<double, std::micro>(r->t_admit - r->t_submit).count() : 0.0;
        out->latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
    }
    s->tickets.erase(it);
    delete r;
    if (status != GHCM_OK) return fail(status, "request %llu: %s", (unsigned long long)ticket,
                                       s->fatal ? s->fatal_msg.c_str() : ghcm_status_string(status));
    return GHCM_OK;
}

extern "C" int ghcm_serve_get_stats(ghcm_server* s, ghcm_serve_stats* out) {
    if (!s || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->pending = s->pending.size();
    out->bytes_in_use = s->arena_alloc.used();
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
    (void)hipSetDevice(s->ordinal);
    s->release_all();
    const int fatal = s->fatal;
    const std::string msg = s->fatal_msg;
    for (auto& kv : s->tickets) delete kv.second;
    delete s;
    if (fatal) return fail(fatal, "server failed: %s", msg.c_str());
    return GHCM_OK;
}
