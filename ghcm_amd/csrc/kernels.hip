// kernels.hip -- gfx950 kernels of the gHCM hot path.
//
// One wave64 work-group solves one tile of TW x TH cells.  The tile and a one-cell
// halo ring (corners included) are staged from HBM into LDS, relaxed there with the
// Godunov upwind update, and the improved cells are written back with plain stores:
// a tile is solved by at most one work-group per wave and no kernel writes another
// tile's cells, so each cell has a single writer per launch.  Neighbours that read
// this tile's edge concurrently see the old or the new value; both are upper bounds,
// and a neighbour that must see the new value is re-queued for the next wave.
//
// Two in-tile styles, as in the reference:
//   MODE_SYSTOLIC  every cell is updated each pass until no cell improves by >= eps,
//                  then the four neighbour-tile border rows are solved and improvements
//                  stored (data/eikonal_ghcm_systolic.cl:8-41, 130-157).
//   MODE_FIM       an open set of cells; converged cells solve their four neighbours,
//                  opening in-tile ones that improve by > eps and flagging the border
//                  when an outside cell would improve by > eps, without storing it
//                  (data/eikonal_ghcm_fim.cl:166-210, 341-372).
// Convergence and border votes are wave ballots; the per-tile min is a butterfly of
// __shfl_xor (replacing the LDS trees of data/util_common.cl:69-141).
//
// Scheduling (device mode): a wave's tiles sit in list[cur]; k_wave_end (one
// work-group) picks the n lowest keys with a two-level radix histogram and k_solve
// solves them.  The rest go to the next list: the boundary moves them itself, except for
// long lists after a solve's first boundary, whose unselected tiles it leaves to the tile
// launch (carry_handed_off).  Activations append neighbours to list[cur^1] once per wave
// (atomicMax on qwave), sharded over kShards counters.
// Host mode takes an explicit job list and writes one status record per job
// (src/eikonal_hcm.cpp:396-420).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "ghcm_dev.h"
#include "kernels.h"
#include "tile_geom.h"

namespace ghcm {

enum { MODE_SYSTOLIC = 0, MODE_FIM = 1 };

// Pointers read out of device memory (the instance records) are generic to the compiler,
// which then emits flat accesses (64-bit lane addresses, both wait counters).  They always
// address global memory; the cast lets it emit global loads and stores with a scalar base.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Bytes per field element of a field kind (FP32 cost or speed, FP16 or BF16 speed).
template <int KIND>
constexpr uint32_t field_bytes() { return KIND <= 1 ? 4u : 2u; }

// Cost of the field element at byte address p.
template <int KIND>
__device__ __forceinline__ float load_cost_at(const char* p) {
    if constexpr (KIND == 0) {
        return *gmem(reinterpret_cast<const float*>(p));
    } else if constexpr (KIND == 1) {
        return 1.0f / *gmem(reinterpret_cast<const float*>(p));
    } else if constexpr (KIND == 2) {
        return 1.0f / static_cast<float>(*gmem(reinterpret_cast<const _Float16*>(p)));
    } else {
        uint32_t b = static_cast<uint32_t>(*gmem(reinterpret_cast<const uint16_t*>(p))) << 16;
        return 1.0f / __uint_as_float(b);
    }
}

// Godunov upwind update for |grad T| = c with h = 1, cancellation-safe form, without the
// final min against the old time.  m = min(a,b) is passed in and d = |a-b|; if c < d the 1-D
// update m + c applies (data/eikonal_common.cl:296-297); otherwise T = m + d/2 +
// sqrt((sqrt2 c - d)(sqrt2 c + d))/2, the root of (T-a)^2 + (T-b)^2 = c^2 (:299-300) with
// the discriminant factored and the large common part m added outside the root.
// Branch-free, without a select: the two-axis form is evaluated at e = min(d, c), as
// m + (e + max(root, e))/2.  For d <= c that is the form itself (root >= d there, exactly;
// the max matters only within an ulp of d = c, where it gives m + d = max(a,b), the exact
// value).  For d > c, e = c and the rounded root never exceeds c (sqrt2 rounds down to the
// float constant), so the max yields c and the result is m + c rounded once: the 1-D update,
// bit for bit.  The result therefore never falls as d crosses c (the 1-D value is not below
// the two-axis values next to it), which the border checks rely on: they predict a
// neighbour's update from a subset of the inputs its own solve sees, and a prediction an ulp
// below what the solve then computes would re-activate the pair without end at eps = 0.
// The discriminant lies in [c^2, 2c^2] on every lane, so the square root (the hardware
// v_sqrt_f32, 1 ulp) never sees a negative argument.  e is a v_med3 against 0 (d and c are
// non-negative); when d is NaN (a = b = +INF) the med3 yields the minimum of the other two
// operands, 0, and the result is m + (a finite value) = +INF.
// The large common part m is added by one final rounding of m + (e + root)/2, the
// parenthesis formed at the scale of c: the result then never rounds below max(a,b) where
// the exact root equals it (c = d), so a cell never drops below a neighbour it was computed
// from.  (Rounding m + e/2 first and adding root/2 after breaks that: two cells can then
// lower each other by an ulp at a time without end.)  The final step is
// fma(e + max(root, e), 1/2, m): halving is exact for normal numbers, so it equals m + (the
// rounded half-sum) with one rounding.
// Every update of a cell and every border prediction of a neighbour's update goes through
// this one function.
__device__ __forceinline__ float godunov_raw(float ax, float ay, float m, float c) {
    const float d = fabsf(ax - ay);                // NaN when both are +INF
    const float e = __builtin_amdgcn_fmed3f(d, c, 0.0f);
    const float s = 1.41421356237309515f * c;
    const float root = __builtin_amdgcn_sqrtf((s - e) * (s + e));
    return __builtin_fmaf(e + __builtin_fmaxf(root, e), 0.5f, m);
}

// The update min'ed with the old time t (:305-308).
__device__ __forceinline__ float godunov(float ax, float ay, float c, float t) {
    return fminf(t, godunov_raw(ax, ay, fminf(ax, ay), c));
}

// Wave reductions: a DPP butterfly inside each row of 16 lanes (quad_perm [1,0,3,2] and
// [2,3,0,1], then row_half_mirror and row_mirror, after which every lane of a row holds its
// row's result), then the four rows' results read into scalar registers and combined there.
// No LDS round trip (a __shfl_xor butterfly is six dependent ds_bpermute).  Every lane must be
// active, as at every call site; the result is wave-uniform.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, CTRL, 0xF, 0xF, false);
}

template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v, Op op) {
    v = op(v, dpp_u32<0xB1>(v));
    v = op(v, dpp_u32<0x4E>(v));
    v = op(v, dpp_u32<0x141>(v));
    v = op(v, dpp_u32<0x140>(v));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return op(op(r0, r1), op(r2, r3));
}

// Minimum of times or keys: non-negative or +INF, so their bit patterns order like the values.
__device__ __forceinline__ float wave_min(float v) {
    return __uint_as_float(wave_reduce_u32(__float_as_uint(v), [](uint32_t x, uint32_t y) { return min(x, y); }));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return wave_reduce_u32(v, [](uint32_t x, uint32_t y) { return x + y; });
}

// Stores to mapped host memory: system scope, relaxed (each word is read on its own).
__device__ __forceinline__ void sys_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_store64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Per-tile counter increment without waiting for a return value (one writer per tile per
// wave, so the atomic only saves the load of a read-modify-write).
__device__ __forceinline__ void stat_add(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Progress record for the host (done flag and the count of wave boundaries executed),
// written by the first work-group of each tile-solve launch: the store to host memory
// then overlaps a tile visit instead of ending k_wave_end.  The wave counters follow every
// 16 boundaries and once the loop is done.
__device__ __forceinline__ void publish_progress(const SolveArgs& a, const GhcmCtrl* ctrl) {
    if (!a.rec || blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t t = ctrl->ticks, done = ctrl->done;
    if ((t & 15u) == 0u || done) {
        sys_store(&a.rec->wave, ctrl->wave);
        sys_store(&a.rec->busy_waves, ctrl->busy_waves);
    }
    sys_store64(&a.rec->progress, ((unsigned long long)done << 32) | t);
}

// Append tile g to list `which` unless it is already there (qwave), in shard g % kShards.
__device__ __forceinline__ void list_append(const SolveArgs& a, uint32_t which, uint32_t g) {
    const uint32_t sh = g % kShards;
    uint32_t slot = atomicAdd(&a.ctrl->count[which][sh * kCounterStride], 1u);
    a.list[((size_t)which * kShards + sh) * a.shard_cap + slot] = g;
}

// Queue tile nb for the next wave unless it is already queued, recording why (border
// side bit) and the smallest candidate time that activated it.  `nxt` is the list that
// collects the next wave's tiles.
__device__ __forceinline__ void activate_np(const SolveArgs& a, uint32_t nb, uint32_t wave,
                                            uint32_t nxt, uint32_t np, uint32_t mask_bits,
                                            float key) {
    atomicMin(&a.key[(size_t)np * a.tiles + nb], __float_as_uint(key));
    atomicOr(&a.mask[(size_t)np * a.tiles + nb], mask_bits);
    uint32_t old = atomicMax(&a.qwave[nb], wave + 2u);
    if (old < wave + 2u) list_append(a, nxt, nb);
}

// The same with the next wave's record the other parity of p (two records).
__device__ __forceinline__ void activate(const SolveArgs& a, uint32_t nb, uint32_t wave,
                                         uint32_t nxt, uint32_t p, uint32_t mask_bits,
                                         float key) {
    activate_np(a, nb, wave, nxt, p ^ 1u, mask_bits, key);
}

// Tiles the wave boundary did not select and left to this launch: each moves its activation
// record to the next parity and is queued for the next wave through activate(), whose qwave
// test keeps it from entering the list twice when this wave's own activations reach it
// first.  Every thread of the grid takes a share; only unselected tiles are touched, and this
// wave's claims and activations use other records, so nothing races.
__device__ __forceinline__ void carry_handed_off(const SolveArgs& a, uint32_t nj, uint32_t ncarry,
                                                 uint32_t wave, uint32_t nxt, uint32_t p) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncarry; i += stride) {
        const uint32_t g = a.jobs[nj + i];
        uint32_t* kp = &a.key[(size_t)p * a.tiles + g];
        uint32_t* mp = &a.mask[(size_t)p * a.tiles + g];
        const uint32_t k = *kp, mb = *mp;
        *kp = kInfBits;
        *mp = 0u;
        activate(a, g, wave, nxt, p, mb, __uint_as_float(k));
    }
}

// Entry i of the current wave list: shard found by a search over the prefix counts.
__device__ __forceinline__ uint32_t list_entry(const SolveArgs& a, const uint32_t* prefix,
                                               uint32_t cur, uint32_t i) {
    uint32_t lo = 0, hi = kShards;           // prefix[lo] <= i < prefix[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid; else hi = mid;
    }
    return a.list[((size_t)cur * kShards + lo) * a.shard_cap + (i - prefix[lo])];
}

// Job j of the current wave: entry j of list[cur] when every queued tile was selected,
// else entry j of the compacted top-n selection.
__device__ __forceinline__ uint32_t job_tile(const SolveArgs& a, const uint32_t* prefix,
                                             uint32_t cur, uint32_t from_list, uint32_t j) {
    return from_list ? list_entry(a, prefix, cur, j) : a.jobs[j];
}

// Claim a job's activation record: its border bits (and open-all flag) for this wave;
// the key and bits are reset for the tile's next activation.
__device__ __forceinline__ uint32_t claim_tile(const SolveArgs& a, uint32_t g, uint32_t p) {
    uint32_t* kp = &a.key[(size_t)p * a.tiles + g];
    uint32_t* mp = &a.mask[(size_t)p * a.tiles + g];
    const uint32_t mb = *mp;
    *kp = kInfBits;
    *mp = 0u;
    return mb;
}

// An instance record as wave-uniform register values.  The record is the same for every
// lane of a work-group; passing each word through readfirstlane puts it in scalar registers,
// so the tile's loads take a scalar base and carry no dependency on the vector load that
// fetched the record (which would otherwise make every conditional staging load wait for
// all loads before it).
__device__ __forceinline__ GhcmInst uniform_inst(const GhcmInst& m) {
    auto u64 = [](const void* p) {
        const uint64_t v = reinterpret_cast<uint64_t>(p);
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        return ((uint64_t)hi << 32) | lo;
    };
    GhcmInst r;
    r.field = reinterpret_cast<const void*>(u64(m.field));
    r.time = reinterpret_cast<float*>(u64(m.time));
    r.w = __builtin_amdgcn_readfirstlane(m.w);
    r.h = __builtin_amdgcn_readfirstlane(m.h);
    r.tiles_x = __builtin_amdgcn_readfirstlane(m.tiles_x);
    r.tiles_y = __builtin_amdgcn_readfirstlane(m.tiles_y);
    r.tile_base = __builtin_amdgcn_readfirstlane(m.tile_base);
    r.d = __builtin_amdgcn_readfirstlane(m.d);
    r.tiles_z = __builtin_amdgcn_readfirstlane(m.tiles_z);
    r.pad = 0u;
    return r;
}

// The record of a single-instance solve, read once per work-group before its job loop.
__device__ __forceinline__ GhcmInst single_inst(const SolveArgs& a) {
    GhcmInst r{};
    if (a.n_inst == 1) r = uniform_inst(a.inst[0]);
    return r;
}

// Stage a PW x PH box of time and cost (the tile plus its halo ring, origin (ox-1, oy-1))
// into LDS; out-of-grid cells read +INF.  Every global load of a thread is issued before
// the first LDS store, so the box costs one memory round trip rather than one per row.
template <int PW, int PH, int NT, int KIND>
__device__ __forceinline__ void stage_tile(const GhcmInst& in, int ox, int oy, int tid,
                                           float* t_s, float* c_s) {
    constexpr int PP = PW * PH;
    constexpr int NL = (PP + NT - 1) / NT;
    const int W = (int)in.w, H = (int)in.h;
    // Addresses are a wave-uniform box origin plus a 32-bit byte offset per cell (the width
    // limit of the C-ABI keeps PH rows of a grid below 2^32 bytes), so every load has a
    // scalar base and one vector offset.
    const int64_t org = (int64_t)(oy - 1) * W + (ox - 1);
    const char* const tb = reinterpret_cast<const char*>(in.time) + org * 4;
    const char* const fb = static_cast<const char*>(in.field) + org * (int64_t)field_bytes<KIND>();
    float tv[NL], cv[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = tid + NT * k;
        const int bx = i % PW, by = i / PW;
        const int gx = ox + bx - 1, gy = oy + by - 1;
        tv[k] = __builtin_inff();
        cv[k] = __builtin_inff();
        if (i < PP && gx >= 0 && gy >= 0 && gx < W && gy < H) {
            const uint32_t off = (uint32_t)by * (uint32_t)W + (uint32_t)bx;
            tv[k] = *gmem(reinterpret_cast<const float*>(tb + off * 4u));
            cv[k] = load_cost_at<KIND>(fb + off * field_bytes<KIND>());
        }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = tid + NT * k;
        if (i < PP) {
            t_s[i] = tv[k];
            c_s[i] = cv[k];
        }
    }
}

// Threads of a sweep work-group: four wave64s, one per sweep order.
constexpr int kSweepThreads = 256;

// Stage a sweep tile: the same PW x PH box as stage_tile (tile plus halo ring, out-of-grid
// cells +INF), read so that no cache line is fetched for a few of its bytes.  The interior is
// TH aligned rows of TW cells of the time and cost fields; the top and bottom halo rows are
// read from the fields (the time rows with their corners); the left and right halo columns
// come from the neighbours' edge columns (a.edge, a.cedge: TH contiguous floats each) instead
// of one line per row.  Every global load is issued before the first LDS store.
template <int TW, int TH, int KIND>
__device__ __forceinline__ void stage_sweep_tile(const SolveArgs& a, const GhcmInst& in, uint32_t g,
                                                 int tx, int ox, int oy, int tid, float* t_s,
                                                 float* c_s, float* told) {
    constexpr int NT = kSweepThreads;
    constexpr int PW = TW + 2, PH = TH + 2, P = TW * TH;
    constexpr int NI = (P + NT - 1) / NT;        // interior cells per thread
    constexpr int NHX = 2 * PW + 2 * TH;         // halo cells: two rows with corners, two columns
    constexpr int NH = (NHX + NT - 1) / NT;
    const int W = (int)in.w, H = (int)in.h;
    // Addresses are a wave-uniform origin plus a 32-bit byte offset per cell (the width limit
    // of the C-ABI keeps TH + 2 rows of a grid below 2^32 bytes).
    const int64_t org = (int64_t)(oy - 1) * W + (ox - 1);
    const char* const tb = reinterpret_cast<const char*>(in.time) + org * 4;
    const char* const fb = static_cast<const char*>(in.field) + org * (int64_t)field_bytes<KIND>();
    const bool has_l = tx > 0, has_r = ox + TW < W;
    // left neighbour's last column and right neighbour's first column (unused at the borders)
    const size_t lo = has_l ? ((size_t)g - 1) * 2 * TH + TH : 0, ro = has_r ? ((size_t)g + 1) * 2 * TH : 0;
    const float* const el = a.edge + lo;
    const float* const er = a.edge + ro;
    const float* const cl = a.cedge + lo;
    const float* const cr = a.cedge + ro;
    float tv[NI], cv[NI], th[NH], ch[NH];
    auto cell_off = [W](int bx, int by) { return (uint32_t)by * (uint32_t)W + (uint32_t)bx; };
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const int i = tid + NT * k;
        const int r = i / TW, c = i % TW;
        tv[k] = __builtin_inff();
        cv[k] = __builtin_inff();
        if (i < P && ox + c < W && oy + r < H) {
            const uint32_t off = cell_off(c + 1, r + 1);
            tv[k] = *gmem(reinterpret_cast<const float*>(tb + off * 4u));
            cv[k] = load_cost_at<KIND>(fb + off * field_bytes<KIND>());
        }
    }
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int h = tid + NT * k;
        th[k] = __builtin_inff();
        ch[k] = __builtin_inff();
        if (h < 2 * PW) {
            // top (box row 0, y = oy - 1) or bottom (box row PH - 1, y = oy + TH) halo row
            const int by = h < PW ? 0 : PH - 1, bx = h % PW;
            const int gx = ox + bx - 1, gy = oy + by - 1;
            if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
                const uint32_t off = cell_off(bx, by);
                th[k] = *gmem(reinterpret_cast<const float*>(tb + off * 4u));
                // the cost box's corners are never read (the border checks use its rows
                // and columns only)
                if (bx > 0 && bx < PW - 1) ch[k] = load_cost_at<KIND>(fb + off * field_bytes<KIND>());
            }
        } else if (h < NHX) {
            const int j = h - 2 * PW, right = j >= TH ? 1 : 0, r = j - right * TH;
            if (oy + r < H && (right ? has_r : has_l)) {
                th[k] = *gmem(right ? er + r : el + r);
                ch[k] = *gmem(right ? cr + r : cl + r);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const int i = tid + NT * k;
        if (i < P) {
            const int q = (i / TW + 1) * PW + (i % TW) + 1;
            t_s[q] = tv[k];
            c_s[q] = cv[k];
        }
        // The staged time of interior cell tid + NT k (+INF outside the grid or the tile):
        // the caller's snapshot, which it would otherwise read back from LDS.
        told[k] = tv[k];
    }
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int h = tid + NT * k;
        int q = -1;
        if (h < 2 * PW) {
            q = (h < PW ? 0 : PH - 1) * PW + h % PW;
        } else if (h < NHX) {
            const int j = h - 2 * PW, right = j >= TH ? 1 : 0, r = j - right * TH;
            q = (r + 1) * PW + (right ? PW - 1 : 0);
        }
        if (q >= 0) {
            t_s[q] = th[k];
            c_s[q] = ch[k];
        }
    }
}

// Edge columns of tile g from its LDS box (after its last round): wave 0 writes the first
// column, wave 1 the last one when the tile is full width (only then has it a right
// neighbour that reads it).  Rows past the grid are never read and not written.
template <int TW, int TH>
__device__ __forceinline__ void write_edges(const SolveArgs& a, uint32_t g, int sx, int sy, int w,
                                            int L, const float* t_s) {
    constexpr int PW = TW + 2;
    if (w == 0) {
        for (int r = L; r < sy; r += 64) *gmem(a.edge + (size_t)g * 2 * TH + r) = t_s[(r + 1) * PW + 1];
    } else if (w == 1 && sx == TW) {
        for (int r = L; r < sy; r += 64) *gmem(a.edge + (size_t)g * 2 * TH + TH + r) = t_s[(r + 1) * PW + TW];
    }
}

template <int TW, int TH, int KIND, int MODE, bool HOST>
__device__ void process_tile(const SolveArgs& a, uint32_t g, uint32_t mbits, uint32_t wave,
                             uint32_t nxt, uint32_t p, uint32_t job_slot,
                             float* t_s, float* c_s, uint8_t* open_s, uint8_t* add_s,
                             uint16_t* list_s) {
    constexpr int PW = TW + 2;
    constexpr int PH = TH + 2;
    constexpr int P = TW * TH;
    constexpr int CPL = (P + 63) / 64;
    const int lane = threadIdx.x;

    const uint32_t ii = (a.n_inst == 1) ? 0u : a.tinst[g];
    const GhcmInst in = uniform_inst(a.inst[ii]);
    int tx, ty;
    tile_xy(g - in.tile_base, in.tiles_x, tx, ty);
    const int ox = tx * TW, oy = ty * TH;
    const int W = (int)in.w, H = (int)in.h;
    const int sx = tile_extent(W, ox, TW), sy = tile_extent(H, oy, TH);

    // ---- stage tile + halo ring into LDS (out of grid -> +INF) ----
    stage_tile<PW, PH, 64, KIND>(in, ox, oy, lane, t_s, c_s);
    __syncthreads();

    float told[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        int i = lane + 64 * k;
        told[k] = (i < P) ? t_s[(i / TW + 1) * PW + (i % TW) + 1] : __builtin_inff();
    }

    const float eps = a.eps;
    const uint32_t hard_cap = 4u * P + 64u;
    const uint32_t cap = (a.iter_cap == 0u || a.iter_cap > hard_cap) ? hard_cap : a.iter_cap;
    uint32_t passes = 0;
    bool unfinished = false;

    uint32_t flags = 0;                       // border dirty bits (per lane, OR-reduced)
    float kc[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
    const bool has_nb[4] = {tx > 0, ty > 0, tx + 1 < (int)in.tiles_x, ty + 1 < (int)in.tiles_y};

    // Halo-cell solve for side s at padded index q: the neighbour further out lies
    // beyond the staged ring and reads as +INF (data/eikonal_common.cl:201-204).
    auto halo_solve = [&](int s, int q) -> float {
        float ax, ay;
        if (s == 0)      { ax = t_s[q + 1];  ay = fminf(t_s[q - PW], t_s[q + PW]); }
        else if (s == 2) { ax = t_s[q - 1];  ay = fminf(t_s[q - PW], t_s[q + PW]); }
        else if (s == 1) { ay = t_s[q + PW]; ax = fminf(t_s[q - 1], t_s[q + 1]); }
        else             { ay = t_s[q - PW]; ax = fminf(t_s[q - 1], t_s[q + 1]); }
        return godunov(ax, ay, c_s[q], t_s[q]);
    };
    auto cell_solve = [&](int q) -> float {
        float ax = fminf(t_s[q - 1], t_s[q + 1]);
        float ay = fminf(t_s[q - PW], t_s[q + PW]);
        return godunov(ax, ay, c_s[q], t_s[q]);
    };

    if constexpr (MODE == MODE_SYSTOLIC) {
        // data/eikonal_ghcm_systolic.cl:130-153: sweep all cells until converged.
        bool go = true;
        while (go) {
            bool changed = false;
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                int i = lane + 64 * k;
                if (i < P) {
                    int q = (i / TW + 1) * PW + (i % TW) + 1;
                    float t0 = t_s[q];
                    float tn = cell_solve(q);
                    if (tn < t0) {
                        t_s[q] = tn;
                        if (t0 - tn >= eps) changed = true;
                    }
                }
            }
            __syncthreads();
            ++passes;
            bool any = __ballot(changed) != 0ull;
            if (any && passes >= cap) unfinished = true;
            go = any && passes < cap;
        }
        // check_border (systolic.cl:8-41): solve the neighbour's adjacent row/column and
        // flag the side if a cell would improve by >= eps.  The improved values are not
        // stored into the neighbour's cells: the flagged neighbour recomputes them from
        // this tile's committed edge on its next visit, so each cell has one writer per
        // wave and commits can be plain stores.
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (!has_nb[s]) continue;
            const int len = (s & 1) ? sx : sy;
            for (int j = lane; j < len; j += 64) {
                int bx, by;
                if (s == 0)      { bx = 0;      by = j + 1; }
                else if (s == 2) { bx = sx + 1; by = j + 1; }
                else if (s == 1) { bx = j + 1;  by = 0; }
                else             { bx = j + 1;  by = sy + 1; }
                int q = by * PW + bx;
                float t0 = t_s[q];
                float tn = halo_solve(s, q);
                if (tn < t0 && t0 - tn >= eps) {
                    flags |= 1u << s;
                    kc[s] = fminf(kc[s], tn);
                }
            }
        }
    } else {
        // FIM body (data/eikonal_ghcm_fim.cl:290-372) with a byte open set in LDS, compacted
        // into a list of open cells at the start of every pass (the reference compacts its
        // open list too, data/eikonal_ghcm_fim.cl:51-134): both phases of a pass touch the open
        // cells only, 64 per iteration, in cell-index order.
        for (int i = lane; i < P; i += 64) {
            int x = i % TW, y = i / TW;
            bool in_grid = x < sx && y < sy;
            bool o = (mbits & kMaskOpenAll) != 0u;
            o = o || ((mbits & 1u) && x == 0) || ((mbits & 4u) && x == sx - 1) ||
                ((mbits & 2u) && y == 0) || ((mbits & 8u) && y == sy - 1);
            open_s[i] = (o && in_grid) ? 1 : 0;
            add_s[i] = 0;
        }
        __syncthreads();
        const uint64_t below = (1ull << lane) - 1ull;
        for (;;) {
            // Compaction: each chunk of 64 cells appends its open ones (ballot, popcount).
            uint32_t n_open = 0;
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const int i = lane + 64 * k;
                const bool o = i < P && open_s[i] != 0;
                const uint64_t m = __ballot(o);
                if (o) list_s[n_open + (uint32_t)__popcll(m & below)] = (uint16_t)i;
                n_open += (uint32_t)__popcll(m);
            }
            if (n_open == 0u) break;
            __syncthreads();
            bool conv[CPL];
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const uint32_t e = (uint32_t)(lane + 64 * k);
                conv[k] = false;
                if (e < n_open) {
                    const int i = list_s[e];
                    int q = (i / TW + 1) * PW + (i % TW) + 1;
                    float t0 = t_s[q];
                    float tn = cell_solve(q);
                    bool improved = false;
                    if (tn < t0) {
                        t_s[q] = tn;
                        improved = (t0 - tn >= eps);
                    }
                    conv[k] = !improved;          // converged iff diff > -eps
                }
            }
            if (passes >= cap) { unfinished = true; break; }
            ++passes;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const uint32_t e = (uint32_t)(lane + 64 * k);
                if (e < n_open && conv[k]) {
                    const int i = list_s[e];
                    const int x = i % TW, y = i / TW;
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        const int nx = x + (d == 0 ? -1 : d == 2 ? 1 : 0);
                        const int ny = y + (d == 1 ? -1 : d == 3 ? 1 : 0);
                        if (nx >= 0 && ny >= 0 && nx < sx && ny < sy) {
                            const int ni = ny * TW + nx;
                            if (open_s[ni]) continue;
                            const int qn = (ny + 1) * PW + nx + 1;
                            float t0 = t_s[qn];
                            float tn = cell_solve(qn);
                            if (tn < t0) {
                                atomicMin(reinterpret_cast<unsigned int*>(t_s) + qn, __float_as_uint(tn));
                                if (t0 - tn > eps) add_s[ni] = 1;
                            }
                        } else if (has_nb[d] && ((d & 1) ? (nx >= 0 && nx < sx) : (ny >= 0 && ny < sy))) {
                            // outside the tile: the cell belongs to neighbour tile d.
                            const int qn = (ny + 1) * PW + nx + 1;
                            float t0 = t_s[qn];
                            float tn = halo_solve(d, qn);
                            if (tn < t0 && t0 - tn > eps) {
                                flags |= 1u << d;
                                kc[d] = fminf(kc[d], tn);
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // Converged open cells leave the set; cells added in this pass join it (they were
            // not open, so the two writes never meet a cell twice).
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const uint32_t e = (uint32_t)(lane + 64 * k);
                if (e < n_open && conv[k]) open_s[list_s[e]] = 0;
            }
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const int i = lane + 64 * k;
                if (i < P && add_s[i]) {
                    open_s[i] = 1;
                    add_s[i] = 0;
                }
            }
            __syncthreads();
        }
    }

    __syncthreads();
    // ---- commit improved cells (plain stores: this work-group is the only writer of
    // the tile's cells in this wave), tile min ----
    char* const cb = reinterpret_cast<char*>(in.time) + ((int64_t)oy * W + ox) * 4;
    float tmin = __builtin_inff();
    uint32_t improved = 0;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        int i = lane + 64 * k;
        if (i < P) {
            int x = i % TW, y = i / TW;
            if (x < sx && y < sy) {
                float tn = t_s[(y + 1) * PW + x + 1];
                tmin = fminf(tmin, tn);
                if (tn < told[k]) {
                    *gmem(reinterpret_cast<float*>(cb + ((uint32_t)y * (uint32_t)W + (uint32_t)x) * 4u)) = tn;
                    ++improved;
                }
            }
        }
    }
    tmin = wave_min(tmin);
    improved = wave_sum(improved);
    uint32_t side_flags = 0;
    float side_key[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (__ballot((flags >> s) & 1u) != 0ull) side_flags |= 1u << s;
        side_key[s] = wave_min(kc[s]);
    }

    if constexpr (HOST) {
        if (lane == 0) {
            HostStatus st;
            st.solves = passes;
            st.cells_updated = improved;
            float sm = tmin;
            for (int s = 0; s < 4; ++s) sm = fminf(sm, side_key[s]);
            st.smallest_time = sm;
            for (int s = 0; s < 4; ++s) st.border[s] = (side_flags >> s) & 1u;
            st.requeue = unfinished ? 1u : 0u;
            a.status[job_slot] = st;
            a.tmin[g] = tmin;
            stat_add(&a.tstat[g], 1u);
            stat_add(&a.tstat[(size_t)a.tiles + g], passes);
            stat_add(&a.tstat[2 * (size_t)a.tiles + g], improved);
        }
    } else {
        // Lanes 0-3 queue the dirty sides' neighbours and lane 4 the tile itself when the pass
        // cap stopped it early, in one call, so their atomics' round trips overlap.
        bool act = false;
        uint32_t nb = g, bits = kMaskOpenAll;
        float key = tmin;
        if (lane < 4 && ((side_flags >> lane) & 1u)) {
            act = true;
            nb = tile_neighbour(in.tile_base, in.tiles_x, in.tiles_y, tx, ty, lane);
            bits = border_bit_across(lane);
            key = side_key[lane & 3];
        } else if (lane == 4) {
            act = unfinished;
        }
        if (act) activate(a, nb, wave, nxt, p, bits, key);
        if (lane == 5) {
            a.tmin[g] = tmin;
            stat_add(&a.tstat[g], 1u);
            stat_add(&a.tstat[(size_t)a.tiles + g], passes);
            stat_add(&a.tstat[2 * (size_t)a.tiles + g], improved);
        }
    }
}

template <int TW, int TH, int KIND, int MODE, bool HOST>
__global__ __launch_bounds__(64) void k_solve(SolveArgs a) {
    constexpr int PP = (TW + 2) * (TH + 2);
    constexpr int P = TW * TH;
    __shared__ float t_s[PP];
    __shared__ float c_s[PP];
    __shared__ uint8_t open_s[MODE == MODE_FIM ? P : 1];
    __shared__ uint8_t add_s[MODE == MODE_FIM ? P : 1];
    __shared__ uint16_t list_s[MODE == MODE_FIM ? P : 1];
    const int lane = threadIdx.x;

    if constexpr (HOST) {
        const uint32_t j = blockIdx.x;
        if (j >= a.n_host_jobs) return;
        const HostJob job = a.host_jobs[j];
        // A graph-replayed wave launches n work-groups; its unused job slots hold kNoTile.
        if (job.id == kNoTile) return;
        process_tile<TW, TH, KIND, MODE, true>(a, job.id, job.border, 0u, 0u, 0u, j,
                                               t_s, c_s, open_s, add_s, list_s);
    } else {
        __shared__ uint32_t s_prefix[kShards + 1];
        const GhcmCtrl* ctrl = a.ctrl;
        publish_progress(a, ctrl);
        // The control words and the shard prefix are read in one round trip; the done check
        // follows.
        const uint32_t done = ctrl->done;
        const uint32_t nj = ctrl->n_jobs;
        const uint32_t wave = ctrl->wave;
        const uint32_t cur = ctrl->cur;
        const uint32_t from_list = ctrl->jobs_from_list;
        const uint32_t ncarry = ctrl->n_carry;
        static_assert(kShards < 64, "one lane per prefix entry");
        // Every lane reads a valid word (no branch around the load).
        const uint32_t pv = ctrl->prefix[min(lane, (int)kShards)];
        // Work-groups past the jobs still take their share of the handed-off carry.
        // A compiler barrier: every load above is issued before the done test (without it the
        // short-circuit test lets the compiler sink the later loads below it, one memory round
        // trip each).
        __asm__ volatile("" ::: "memory");
        if (done || (blockIdx.x >= nj && ncarry == 0u)) return;
        const uint32_t p = wave & 1u;
        if (lane <= (int)kShards) s_prefix[lane] = pv;
        __syncthreads();
        for (uint32_t j = blockIdx.x; j < nj; j += gridDim.x) {
            const uint32_t g = job_tile(a, s_prefix, cur, from_list, j);
            uint32_t mb = 0;
            if (lane == 0) mb = claim_tile(a, g, p);
            mb = __shfl(mb, 0);
            process_tile<TW, TH, KIND, MODE, false>(a, g, mb, wave, cur ^ 1u, p, j,
                                                    t_s, c_s, open_s, add_s, list_s);
            __syncthreads();
        }
        if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
    }
}

// ---------------------------------------------------------------------------------------
// Sweep tile solver: four wave64s per work-group relax one TW x TH tile (TW <= 64) with
// the four Gauss-Seidel orders of fast sweeping (Zhao 2005, the method
// data/eikonal_common.cl:227 cites), one order per wave, concurrently on the same LDS tile.
// Lane L owns column x (x = L for sweeps towards +x, TW-1-L towards -x) and handles row
// index (s-L) mod TH at step s (a wrapped, skewed anti-diagonal front), so the upwind column
// neighbour of the current row was produced by lane L-1 one step earlier: it arrives as a
// wave_shr:1 DPP operand of the minimum that merges it with the LDS read, and the upwind row
// neighbour is the lane's own previous result.  Every neighbour is also read from LDS (any
// value there is an upper bound, possibly already lowered by another wave), and results are
// merged into the tile with an LDS atomic minimum.  A round is the four sweeps; rounds repeat
// until no cell of the tile improves by >= eps or the round cap is reached (the tile then
// re-queues itself).
// ---------------------------------------------------------------------------------------
constexpr uint32_t kSweepHardRounds = 64;

// min(u, lane L-1's v) on time bits, with lane 0 (no source lane) taking u.  The DPP move's
// `old` is the minimum's identity (all ones) and its bound control is off, so the compiler
// folds the move into the minimum as one v_min_u32 with a wave_shr:1 DPP operand, whose
// lane 0 keeps u.
__device__ __forceinline__ float umin_from_lower_lane(float u, float v) {
    const uint32_t w = (uint32_t)__builtin_amdgcn_update_dpp(-1, __float_as_int(v), 0x138, 0xF, 0xF, false);
    return __uint_as_float(min(__float_as_uint(u), w));
}

// LDS words of a sweep tile: time box, a sink word, cost box.
template <int TW, int TH>
constexpr int sweep_lds_words() { return 2 * (TW + 2) * (TH + 2) + 1; }

// One round of one sweep order over the LDS tile (time box at shb, cost box CB words after it),
// run by one wave64.  DX > 0: lane L owns column x = L, sweeping towards +x; DX < 0: column
// TW-1-L, towards -x.  DOWN: rows ascending.  Wrapped skew: at step s lane L relaxes pass
// position (s - L) mod TH of its column (row p of a downward pass, TH-1-p of an upward one),
// so lane L reaches the first row of its ordered pass at step L, one step after lane L-1, and
// the lanes' ordered passes form one anti-diagonal front; TW + TH - 1 steps give every column
// one complete ordered pass, and before and after it a lane relaxes cells of the neighbouring
// passes instead of idling (every value involved is an upper bound).
//
// A step reads only the cell's two upwind neighbours and its cost: the sweep towards +x and +y
// updates a cell from its -x and -y neighbours.  The other two sweeps' updates cover the +x and
// +y neighbours, and the minimum of the four one-sided updates is the two-sided Godunov update
// (it is monotone in both axis values), so the four sweeps together have the same fixed point
// as sweeps that read all four neighbours.  The upwind values just computed arrive through
// `prev` (this lane's previous cell, the upwind row) and a wave_shr:1 DPP move of lane L-1's
// `prev` (the upwind column); each is merged with this lane's own LDS read of that cell, so
// improvements made by the other three sweeps are used as soon as they land.  Lane 0 has no
// source lane: its upwind column is the halo read from LDS alone.
// Results are merged into the tile with an LDS atomic minimum (no return value waited for).
template <int TW, int TH, int DX, bool DOWN>
__device__ __forceinline__ void sweep_round(char* const shb, const int L) {
    constexpr int PW = TW + 2;
    constexpr int PH = TH + 2;
    constexpr int CB = PW * PH + 1;
    static_assert((TH & (TH - 1)) == 0, "TH must be a power of two");
    const bool lane_on = L < TW;
    // Lanes past the tile width (TW < 64) shadow the last column and never store.
    const int xc = lane_on ? (DX > 0 ? L : TW - 1 - L) : (DX > 0 ? TW - 1 : 0);
    // Word offsets from the corner (r, xc) of the 3 x 3 box around the cell, which sits at
    // padded position (r + 1, xc + 1): every read and the store are immediate offsets from one
    // per-step address.
    constexpr int kUpRow = DOWN ? 1 : 2 * PW + 1;
    constexpr int kUpCol = DX > 0 ? PW : PW + 2;
    constexpr int kCell = PW + 1;
    constexpr int kCost = CB + PW + 1;
    constexpr int b_step = DOWN ? 4 * PW : -4 * PW;
    constexpr int r_first = DOWN ? 0 : TH - 1;
    const int p0 = (-L) & (TH - 1);
    const int r_start = DOWN ? p0 : TH - 1 - p0;
    const int b_first = r_first * (4 * PW) + 4 * xc;
    const int b_start = r_start * (4 * PW) + 4 * xc;
    // Lane L is at the first row of its pass at the steps s with s = L (mod TH): a scalar
    // one-hot pattern shifted by s & (TH-1) marks those lanes, and the row's byte offset
    // advances by one row per step, back to the first row exactly there (the wrap of the row
    // index coincides with the pass start).
    constexpr uint64_t kFirstPat = TH >= 64 ? 1ull : ~0ull / ((1ull << (TH & 63)) - 1ull);
    constexpr int S = TW + TH - 1;
    constexpr int SE = (S + 1) & ~1;
    int bb = b_start - b_step;
    uint32_t jl = 0;
    struct Nb {
        float ur, uc, cc;
        int b;
        bool first;
    };
    auto at = [&](int b, int word) -> float { return *reinterpret_cast<const float*>(shb + b + 4 * word); };
    // LDS reads of step s+1 are issued before step s is computed (two register sets): their
    // cells are upper bounds whatever the interleaving, and the fresh upwind values arrive
    // through DPP and `prev`, so the dependent chain per step is DPP + update.
    auto load = [&](Nb& o) {
        o.first = __builtin_amdgcn_inverse_ballot_w64(kFirstPat << (jl & (TH - 1)));
        ++jl;
        bb = o.first ? b_first : bb + b_step;
        o.b = bb;
        o.ur = at(o.b, kUpRow);
        o.uc = at(o.b, kUpCol);
        o.cc = at(o.b, kCost);
    };
    // Times are non-negative or +INF, so their bit patterns order like the values: the axis
    // minima and the merge use integer minimum (no NaN canonicalisation needed).
    auto umin = [](float u, float v) {
        return __uint_as_float(min(__float_as_uint(u), __float_as_uint(v)));
    };
    float prev = __builtin_inff();
    auto step = [&](const Nb& n) {
        // Lane L-1's fresh value of the upwind column cell.
        const float ax = umin_from_lower_lane(n.uc, prev);
        const float ay = umin(n.ur, n.first ? __builtin_inff() : prev);
        const float t = godunov_raw(ax, ay, umin(ax, ay), n.cc);
        if (TW == 64 || lane_on)
            __hip_atomic_fetch_min(reinterpret_cast<uint32_t*>(shb + n.b + 4 * kCell),
                                   __float_as_uint(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        // The raw result is handed on, not min(t, old): both consumers take the minimum with
        // their own LDS read of the cell.
        prev = t;
    };
    Nb A, B;
    load(A);
#pragma unroll 2
    for (int st = 0; st < SE; st += 2) {
        load(B);
        __builtin_amdgcn_sched_barrier(0);
        step(A);
        load(A);
        __builtin_amdgcn_sched_barrier(0);
        step(B);
    }
}

template <int TW, int TH, int KIND, bool HOST>
__device__ void sweep_tile(const SolveArgs& a, uint32_t g, uint32_t wave, uint32_t nxt,
                           uint32_t np, uint32_t job_slot, float* sh,
                           uint32_t* s_u32, float* s_f32, GhcmInst in_single) {
    constexpr int PW = TW + 2;
    constexpr int PH = TH + 2;
    constexpr int PP = PW * PH;
    constexpr int CB = PP + 1;                     // cost box offset: t_s | sink | c_s
    float* const t_s = sh;
    float* const c_s = sh + CB;
    constexpr int P = TW * TH;
    constexpr int CPT = (P + kSweepThreads - 1) / kSweepThreads;
    static_assert(TW <= 64, "one lane per column");
    const int tid = threadIdx.x;
    const int w = tid >> 6, L = tid & 63;

    // One instance: its record was read before the job loop (one dependent load fewer).
    // It comes by value and is copied, not chosen by reference: a choice between two
    // addresses would keep it in scratch behind flat loads.
    GhcmInst in = in_single;
    if (a.n_inst != 1) in = uniform_inst(a.inst[a.tinst[g]]);
    int tx, ty;
    tile_xy(g - in.tile_base, in.tiles_x, tx, ty);
    const int ox = tx * TW, oy = ty * TH;
    const int W = (int)in.w, H = (int)in.h;
    const int sx = tile_extent(W, ox, TW), sy = tile_extent(H, oy, TH);

    // told[k]: the staged time of cell tid + kSweepThreads k (staging maps cells to threads as
    // the commit does), the snapshot the commit compares with.
    float told[CPT];
    stage_sweep_tile<TW, TH, KIND>(a, in, g, tx, ox, oy, tid, t_s, c_s, told);
    __syncthreads();

    const float eps = a.eps;
    const uint32_t cap = (a.iter_cap == 0u || a.iter_cap > kSweepHardRounds) ? kSweepHardRounds : a.iter_cap;
    char* const shb = reinterpret_cast<char*>(sh);
    uint32_t rounds = 0;
    bool unfinished = false;
    // Convergence: a round counts as changing the tile when some cell ends it at least eps
    // (strictly: an improvement of 0 never counts, so eps = 0 terminates) below where it began;
    // each thread compares its cells with a snapshot taken at the round's start.
    float snap[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) snap[k] = told[k];
    for (;;) {
        // Wave w runs sweep order w (the branch is wave-uniform).
        switch (w) {
            case 0: sweep_round<TW, TH, 1, true>(shb, L); break;
            case 1: sweep_round<TW, TH, -1, true>(shb, L); break;
            case 2: sweep_round<TW, TH, 1, false>(shb, L); break;
            default: sweep_round<TW, TH, -1, false>(shb, L); break;
        }
        ++rounds;
        __syncthreads();
        // With one round per visit its change test is made in the commit pass below, which
        // reads the same cells (the round began at the staged times, `told`).
        if (cap == 1u) break;
        bool chg = false;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int i = tid + kSweepThreads * k;
            if (i < P) {
                const float now = t_s[(i / TW + 1) * PW + (i % TW) + 1];
                const float d = snap[k] - now;      // NaN (INF - INF) counts as no change
                if (d > 0.0f && d >= eps) chg = true;
                snap[k] = now;
            }
        }
        const bool any = __ballot(chg) != 0ull;
        if (L == 0) s_u32[w] = any ? 1u : 0u;
        __syncthreads();
        // No third barrier: a wave writes s_u32 again only after the next round's first
        // barrier, which every wave reaches after reading it here.
        const bool more = (s_u32[0] | s_u32[1] | s_u32[2] | s_u32[3]) != 0u;
        if (!more) break;
        if (rounds >= cap) { unfinished = true; break; }
    }
    // One round per visit: whether it changed the tile (then the tile re-queues itself) is
    // tested against `told` in the commit pass (cells outside the grid are +INF and never
    // change).
    const bool one_round = cap == 1u;
    // Commit improved cells first (plain stores: one writer per tile per wave), so the stores
    // drain while the border check and the activations run; the tile min is reduced after.
    char* const cb = reinterpret_cast<char*>(in.time) + ((int64_t)oy * W + ox) * 4;
    float tmin = __builtin_inff();
    uint32_t improved = 0;
    bool last_chg = false;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int i = tid + kSweepThreads * k;
        if (i < P) {
            const int cx = i % TW, cy = i / TW;
            if (cx < sx && cy < sy) {
                const float tn = t_s[(cy + 1) * PW + cx + 1];
                tmin = fminf(tmin, tn);
                if (tn < told[k]) {
                    *gmem(reinterpret_cast<float*>(cb + ((uint32_t)cy * (uint32_t)W + (uint32_t)cx) * 4u)) = tn;
                    ++improved;
                    if (told[k] - tn >= eps) last_chg = true;
                }
            }
        }
    }
    {
        const bool wave_chg = __ballot(one_round && last_chg) != 0ull;
        if (L == 0) s_u32[12 + w] = wave_chg ? 1u : 0u;
    }
    write_edges<TW, TH>(a, g, sx, sy, w, L, t_s);
    // Border check, wave w takes side w: would the neighbour's adjacent cell improve by
    // >= eps from this tile's edge (data/eikonal_ghcm_systolic.cl:8-41)?
    {
        const int side = w;
        const bool has_nb = side == 0 ? tx > 0 : side == 1 ? ty > 0
                          : side == 2 ? tx + 1 < (int)in.tiles_x : ty + 1 < (int)in.tiles_y;
        bool flag = false;
        float kc = __builtin_inff();
        if (has_nb) {
            const int len = (side & 1) ? sx : sy;
            for (int j = L; j < len; j += 64) {
                int bx, by;
                if (side == 0)      { bx = 0;      by = j + 1; }
                else if (side == 2) { bx = sx + 1; by = j + 1; }
                else if (side == 1) { bx = j + 1;  by = 0; }
                else                { bx = j + 1;  by = sy + 1; }
                const int q = by * PW + bx;
                float ax, ay;
                if (side == 0)      { ax = t_s[q + 1];  ay = fminf(t_s[q - PW], t_s[q + PW]); }
                else if (side == 2) { ax = t_s[q - 1];  ay = fminf(t_s[q - PW], t_s[q + PW]); }
                else if (side == 1) { ay = t_s[q + PW]; ax = fminf(t_s[q - 1], t_s[q + 1]); }
                else                { ay = t_s[q - PW]; ax = fminf(t_s[q - 1], t_s[q + 1]); }
                const float t0 = t_s[q];
                const float tn = godunov(ax, ay, c_s[q], t0);
                if (tn < t0 && t0 - tn >= eps) {
                    flag = true;
                    kc = fminf(kc, tn);
                }
            }
        }
        const bool f = __ballot(flag) != 0ull;
        kc = wave_min(kc);
        if (L == 0) {
            s_u32[4 + side] = f ? 1u : 0u;
            s_f32[side] = kc;
        }
    }
    // Per-wave tile minimum and improved count, then one barrier after which every
    // activation is issued at once: the dirty sides' neighbours (lanes 0-3 of wave 0) and,
    // when the round cap stopped the tile before it converged, the tile itself (wave 1).  Each
    // activation is two dependent atomics; issued side by side their round trips overlap
    // instead of following one another at the end of the visit.  The neighbours run in a
    // later wave, after this kernel's stores.
    tmin = wave_min(tmin);
    improved = wave_sum(improved);
    if (L == 0) {
        s_f32[4 + w] = tmin;
        s_u32[8 + w] = improved;
    }
    __syncthreads();
    const float tm = fminf(fminf(s_f32[4], s_f32[5]), fminf(s_f32[6], s_f32[7]));
    if (one_round) unfinished = (s_u32[12] | s_u32[13] | s_u32[14] | s_u32[15]) != 0u;
    if constexpr (!HOST) {
        if (tid < 4 && s_u32[4 + tid]) {
            const int sd = tid;
            const uint32_t nb = tile_neighbour(in.tile_base, in.tiles_x, in.tiles_y, tx, ty, sd);
            activate_np(a, nb, wave, nxt, np, border_bit_across(sd), s_f32[sd]);
        }
        if (tid == 64 && unfinished) activate_np(a, g, wave, nxt, np, kMaskOpenAll, tm);
    }
    if (tid == 128) {
        const uint32_t imp = s_u32[8] + s_u32[9] + s_u32[10] + s_u32[11];
        a.tmin[g] = tm;
        stat_add(&a.tstat[g], 1u);
        stat_add(&a.tstat[(size_t)a.tiles + g], rounds);
        stat_add(&a.tstat[2 * (size_t)a.tiles + g], imp);
        if constexpr (HOST) {
            HostStatus st;
            st.solves = rounds;
            st.cells_updated = imp;
            float sm = tm;
            for (int sd = 0; sd < 4; ++sd) {
                st.border[sd] = s_u32[4 + sd];
                if (s_u32[4 + sd]) sm = fminf(sm, s_f32[sd]);
            }
            st.smallest_time = sm;
            st.requeue = unfinished ? 1u : 0u;
            a.status[job_slot] = st;
        }
    }
}

// WPE: minimum waves per SIMD the register allocation must allow (1: no constraint).  The
// 4-wave form (at most 128 VGPRs, four work-groups per CU) serves waves with many more tiles
// than CUs (batches, serving), where occupancy hides latency; the 1-wave form keeps the
// shorter single-tile latency that a plain solve's waves of a few hundred tiles are bound by.
template <int TW, int TH, int KIND, bool HOST, int WPE = 1>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_solve_sweep(SolveArgs a) {
    __shared__ float sh[sweep_lds_words<TW, TH>()];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[8];
    const int tid = threadIdx.x;
    // Single-instance solves: the instance record is read once, next to the control block.
    const GhcmInst in0 = single_inst(a);
    if constexpr (HOST) {
        const uint32_t j = blockIdx.x;
        if (j >= a.n_host_jobs) return;
        const uint32_t id = a.host_jobs[j].id;
        if (id == kNoTile) return;          // an unused slot of a graph-replayed wave
        sweep_tile<TW, TH, KIND, true>(a, id, 0u, 0u, 0u, j, sh, s_u32, s_f32, in0);
    } else {
        const GhcmCtrl* ctrl = a.ctrl;
        publish_progress(a, ctrl);
        // The control words and the shard prefix are read in one round trip; the done check
        // follows.  Every wave keeps the prefix in its lanes 0..kShards-1.
        const int lane = tid & 63;
        const uint32_t done = ctrl->done;
        const uint32_t nj = ctrl->n_jobs;
        const uint32_t wave = ctrl->wave;
        const uint32_t cur = ctrl->cur;
        const uint32_t from_list = ctrl->jobs_from_list;
        const uint32_t ncarry = ctrl->n_carry;
        // Every lane reads a valid word (no branch around the load).
        const uint32_t pr = ctrl->prefix[lane & (int)(kShards - 1)];
        const uint32_t pv = lane < (int)kShards ? pr : 0xFFFFFFFFu;
        // Work-groups past the jobs still take their share of the handed-off carry.
        // A compiler barrier: every load above is issued before the done test (without it the
        // short-circuit test lets the compiler sink the later loads below it, one memory round
        // trip each).
        __asm__ volatile("" ::: "memory");
        if (done || (blockIdx.x >= nj && ncarry == 0u)) return;
        const uint32_t p = wave & 1u;
        for (uint32_t j = blockIdx.x; j < nj; j += gridDim.x) {
            uint32_t g;
            if (from_list) {
                // Entry j of list[cur]: its shard is the last one whose prefix is <= j (the
                // prefix is non-decreasing and starts at 0), found with one ballot.
                const uint32_t sh = (uint32_t)__popcll(__ballot(pv <= j)) - 1u;
                const uint32_t base = __builtin_amdgcn_readlane(pv, (int)sh);
                g = a.list[((size_t)cur * kShards + sh) * a.shard_cap + (j - base)];
            } else {
                g = a.jobs[j];
            }
            if (tid == 0) claim_tile(a, g, p);
            sweep_tile<TW, TH, KIND, false>(a, g, wave, cur ^ 1u, p ^ 1u, j, sh, s_u32, s_f32, in0);
            __syncthreads();
        }
        if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
    }
}

// Inclusive sums by DPP, without a cross-lane memory round trip: row shifts by 1, 2, 4 and 8
// inside each row of 16 lanes (lanes without a source read 0, the identity), then lane 15 of
// rows 0 and 2 added to rows 1 and 3 (row_bcast:15): each 32-lane half of the wave scanned on
// its own.  The whole wave adds lane 31 to rows 2 and 3 as well (row_bcast:31).  Every lane
// of the wave must be active.
__device__ __forceinline__ uint32_t half_incl_scan_dpp(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    return x;
}

__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x) {
    x = half_incl_scan_dpp(x);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Block-wide exclusive scan of v over NT threads (NT a multiple of 64, at most 1024;
// part[] holds at least 64 + NT/64 words); returns the exclusive prefix of this thread
// and stores the block total in *total.  Each wave scans by DPP, lane 63 posts the wave's
// total and every thread adds up the totals of the waves before its own: two barriers (the
// second leaves part[] free for the next scan).
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* part, uint32_t* total) {
    constexpr int NWV = NT / 64;
    static_assert(NT % 64 == 0 && NWV <= 64, "whole waves");
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t x = wave_incl_scan_dpp(v);
    if (lane == 63) part[wv] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) {
        const uint32_t t = part[w];
        tot += t;
        base += w < wv ? t : 0u;
    }
    *total = tot;
    __syncthreads();
    return base + x - v;
}

// Block-wide exclusive scan over the four wave64s of a 256-thread work-group: each wave scans
// by DPP (wave_incl_scan_dpp) and the wave totals go through part[0..3].  One barrier: a caller
// alternates between two part buffers, so a scan never writes the words another thread may
// still be reading from the one before.
__device__ __forceinline__ uint32_t block_excl_scan4(uint32_t v, uint32_t* part) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t x = wave_incl_scan_dpp(v);
    if (lane == 63) part[wv] = x;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wv; ++w) base += part[w];
    return base + x - v;
}

// ---------------------------------------------------------------------------------------
// Boundary-free waves (launch_mode 4): every tile queued for a wave is solved in that wave,
// so no selection runs and no wave-boundary kernel sits between two tile launches.  Wave w
// (the launch argument) reads list[w % 3], whose shard counts are final because wave w-1
// ended with its launch; every wave64 scans the counts in its lanes and finds a job's shard
// with one ballot.  Activations go to list[(w+1) % 3].  The first work-group empties
// list[(w+2) % 3] (consumed by wave w-1, filled by wave w+1), publishes progress for the
// host, which keeps a few waves enqueued ahead and stops at the first empty one, and keeps
// the wave counters.  The grid (n work-groups) loops over the list, so a wave of any length
// is solved whole.
//
// With SEL (launch_mode 5) the launch also makes the top-n cut itself: when list[w % 3] holds
// more than n tiles, every work-group computes the same selection from the same records --
// the rule of wave_select (two radix levels over key bits [30:19] and [18:7]; keys below the
// threshold bin go, keys inside it go in list order up to the quota) -- so each wave solves
// the tiles k_wave_end would have selected, without a wave-boundary kernel.  Activation
// records then come in three like the lists (record index = list index): wave w reads record
// w % 3, its activations write record (w+1) % 3, and it clears record (w+2) % 3 (read by wave
// w-1, next written by wave w+1), so no record a wave reads changes while it runs.  The
// launch's last kSelCarryGroups work-groups solve nothing: they carry the unselected tiles to
// the next wave (key to record (w+1) % 3, entry to list (w+1) % 3), off the solving groups'
// critical path.  Lists longer than kSelMax, or with more than kSelJobs selected tiles per
// solving group, are solved whole as in launch_mode 4.
// ---------------------------------------------------------------------------------------
constexpr uint32_t kSelPerThread = 16;
constexpr uint32_t kSelMax = kSelPerThread * kSweepThreads;     // 4096 entries
constexpr uint32_t kSelJobs = 8;                                // selected tiles per solving group
constexpr uint32_t kSelCarryGroups = 32;
// LDS words of the selection scratch (histogram, scan partials, shard prefix, broadcast), laid
// over the tile's box before the tile is staged.
constexpr int kSelLdsWords = 4096 + 8 + (kShards + 1) + 4;

template <int TW, int TH, int KIND, int WPE = 1, bool SEL = false>
__global__ __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_solve_sweep_bf(SolveArgs a, uint32_t wave) {
    __shared__ float sh[sweep_lds_words<TW, TH>()];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[8];
    __shared__ uint32_t s_jobs[SEL ? kSelJobs : 1];
    static_assert(!SEL || sweep_lds_words<TW, TH>() >= kSelLdsWords, "selection scratch fits the tile box");
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const GhcmInst in0 = single_inst(a);
    GhcmCtrl* ctrl = a.ctrl;
    const uint32_t cur = wave % 3u;
    const uint32_t nxt = cur == 2u ? 0u : cur + 1u;
    const uint32_t emp = nxt == 2u ? 0u : nxt + 1u;
    static_assert(kShards == 32, "one lane per shard, half a wave");
    // Shard counts of this wave's list: lane l < 32 holds shard l (every lane reads a valid
    // word, so the load needs no branch).
    const uint32_t cnt = ctrl->count[cur][(uint32_t)(lane & (int)(kShards - 1)) * kCounterStride];
    const uint32_t n = SEL ? ctrl->n_select : 0u;
    // Lanes 0-31 scan the counts (lanes 32-63 scan zeros).
    const uint32_t inc = half_incl_scan_dpp(lane < (int)kShards ? cnt : 0u);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)inc, (int)kShards - 1);
    // Exclusive prefix of the shard counts in lanes 0-31; lanes 32-63 never match a job.
    const uint32_t pv = lane < (int)kShards ? inc - cnt : 0xFFFFFFFFu;
    if (blockIdx.x == 0) {
        if (tid < (int)kShards) ctrl->count[emp][tid * kCounterStride] = 0u;
        if (tid == 0) {
            ctrl->wave = wave;
            if (c) {
                __hip_atomic_fetch_add(&ctrl->busy_waves, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                ctrl->done = 1u;
            }
            if (a.rec) sys_store64(&a.rec->progress, ((unsigned long long)(c == 0u) << 32) | (wave + 1u));
        }
    }
    const uint32_t nsolve = SEL ? gridDim.x - kSelCarryGroups : gridDim.x;
    if constexpr (SEL) {
        // Clear record (w+2) % 3, every work-group a share.
        const size_t T = a.tiles;
        for (size_t t = (size_t)blockIdx.x * kSweepThreads + tid; t < T; t += (size_t)gridDim.x * kSweepThreads) {
            a.key[(size_t)emp * T + t] = kInfBits;
            a.mask[(size_t)emp * T + t] = 0u;
        }
    }
    if (c == 0u) return;
    const uint32_t p = wave & 1u;
    // Record the activations of this wave write: the list index with SEL, else the parity.
    const uint32_t np = SEL ? nxt : (p ^ 1u);
    const bool cut = SEL && n > 0u && c > n && c <= kSelMax && n <= kSelJobs * nsolve;
    if (!cut) {
        if (SEL && blockIdx.x >= nsolve) return;
        for (uint32_t j = blockIdx.x; j < c; j += nsolve) {
            // Entry j of list[cur]: its shard is the last one whose prefix is <= j.
            const uint32_t shd = (uint32_t)__popcll(__ballot(pv <= j)) - 1u;
            const uint32_t base = __builtin_amdgcn_readlane(pv, (int)shd);
            const uint32_t g = a.list[((size_t)cur * kShards + shd) * a.shard_cap + (j - base)];
            if (!SEL && tid == 0) claim_tile(a, g, p);
            sweep_tile<TW, TH, KIND, false>(a, g, wave, nxt, np, j, sh, s_u32, s_f32, in0);
            __syncthreads();
        }
        return;
    }
    if constexpr (SEL) {
        // ---- the top-n cut of wave_select, made by every work-group alike ----
        // LDS scratch laid over the tile box: histogram, the keys of every tile (when the
        // grid has at most 4096), scan partials, shard prefix, broadcast words.
        uint32_t* const hist = reinterpret_cast<uint32_t*>(sh);
        constexpr bool kAllKeys = sweep_lds_words<TW, TH>() >= kSelLdsWords + 4096;
        uint32_t* const kall = hist + 4096;
        uint32_t* const part = kall + (kAllKeys ? 4096 : 0);
        uint32_t* const spre = part + 8;
        uint32_t* const smisc = spre + kShards + 1;
        const bool keys_lds = kAllKeys && a.tiles <= 4096u;
        // Lane l < kShards of every wave holds shard l's exclusive prefix.
        if (tid < (int)kShards) spre[tid] = pv;
        if (tid == (int)kShards) spre[kShards] = c;
        __syncthreads();
        // Thread t holds entries [16 t, 16 t + 16) of the list (list order runs over threads).
        const uint32_t i0 = (uint32_t)tid * kSelPerThread;
        const uint32_t* key = a.key + (size_t)cur * a.tiles;
        uint32_t eg[kSelPerThread], ek[kSelPerThread];
        {
            // Every entry's word first (a binary search over the shard prefix in LDS: the last
            // shard with prefix <= i, which is never an empty one), then all list loads (and,
            // for a small grid, the keys of all tiles beside them) unconditionally, indices past
            // the list clamped to its last entry, so no load waits behind a branch or a load.
            uint32_t wi[kSelPerThread];
#pragma unroll
            for (uint32_t u = 0; u < kSelPerThread; ++u) {
                const uint32_t i = min(i0 + u, c - 1u);
                uint32_t sd = 0;
#pragma unroll
                for (uint32_t step = kShards / 2; step >= 1; step >>= 1)
                    sd = spre[sd + step] <= i ? sd + step : sd;
                wi[u] = (cur * kShards + sd) * a.shard_cap + (i - spre[sd]);
            }
            uint32_t kv[kSelPerThread];
#pragma unroll
            for (uint32_t u = 0; u < kSelPerThread; ++u) eg[u] = a.list[wi[u]];
            // The first level's histogram is cleared while the loads are in flight.
            for (int i = tid; i < 4096; i += kSweepThreads) hist[i] = 0u;
            if (keys_lds) {
#pragma unroll
                for (uint32_t u = 0; u < kSelPerThread; ++u) {
                    const uint32_t t = (uint32_t)tid + kSweepThreads * u;
                    kv[u] = key[min(t, a.tiles - 1u)];
                }
#pragma unroll
                for (uint32_t u = 0; u < kSelPerThread; ++u) {
                    const uint32_t t = (uint32_t)tid + kSweepThreads * u;
                    if (t < a.tiles) kall[t] = kv[u];
                }
                __syncthreads();
#pragma unroll
                for (uint32_t u = 0; u < kSelPerThread; ++u) ek[u] = kall[eg[u]];
            } else {
#pragma unroll
                for (uint32_t u = 0; u < kSelPerThread; ++u) ek[u] = key[eg[u]];
                __syncthreads();
            }
#pragma unroll
            for (uint32_t u = 0; u < kSelPerThread; ++u) ek[u] = i0 + u < c ? ek[u] : 0xFFFFFFFFu;
        }
        uint32_t lo_base = 0u, below = 0u;
        for (int level = 0; level < 2; ++level) {
            const int shift = level == 0 ? 19 : 7;
            // A thread's keys are consecutive list entries, and near the front most share a
            // bin: runs of one bin are counted locally and added with one atomic each.
            uint32_t run_b = 0xFFFFFFFFu, run_n = 0u;
#pragma unroll
            for (uint32_t u = 0; u < kSelPerThread; ++u) {
                const uint32_t k = ek[u];
                if (i0 + u < c && (level == 0 || (k >> 19) == (lo_base >> 19))) {
                    const uint32_t b = (k >> shift) & 0xFFFu;
                    if (b != run_b) {
                        if (run_n) atomicAdd(&hist[run_b], run_n);
                        run_b = b;
                        run_n = 0u;
                    }
                    ++run_n;
                }
            }
            if (run_n) atomicAdd(&hist[run_b], run_n);
            __syncthreads();
            constexpr int BPT = 4096 / kSweepThreads;
            uint32_t hb[BPT], mine = 0;
#pragma unroll
            for (int j = 0; j < BPT; ++j) {
                hb[j] = hist[BPT * tid + j];
                mine += hb[j];
            }
            // (The scan's barrier also orders every read of this level's bins before the next
            // level's clear below.)
            const uint32_t excl = block_excl_scan4(mine, part + 4 * level) + below;
            if (excl < n && excl + mine >= n) {           // the crossing bin is one of mine
                uint32_t cum = excl, b = BPT * tid;
                bool found = false;
#pragma unroll
                for (int j = 0; j < BPT; ++j) {
                    if (!found && cum + hb[j] >= n) { b = BPT * tid + j; found = true; }
                    if (!found) cum += hb[j];
                }
                smisc[2 * level] = b;
                smisc[2 * level + 1] = cum;
            }
            if (level == 0)
                for (int i = tid; i < 4096; i += kSweepThreads) hist[i] = 0u;
            __syncthreads();
            lo_base |= smisc[2 * level] << shift;
            below = smisc[2 * level + 1];
        }
        // Keys below lo_base all go (`below` of them); keys in [lo_base, lo_base + 128) go in
        // list order while the quota n - below lasts.  One scan of (below-count << 16 |
        // in-bin count) gives both ranks: the selected entries before this thread's are its
        // below-count prefix plus min(in-bin prefix, quota).
        const uint32_t quota = n - below;
        const uint32_t thr_hi = lo_base + (1u << 7);
        // Job slots start empty (ordered before their assignment by the scan's barriers).
        if (tid < (int)kSelJobs) s_jobs[tid] = kNoTile;
        uint32_t nlo = 0, nib = 0;
#pragma unroll
        for (uint32_t u = 0; u < kSelPerThread; ++u) {
            if (i0 + u >= c) continue;
            if (ek[u] < lo_base) ++nlo;
            else if (ek[u] < thr_hi) ++nib;
        }
        const uint32_t pre = block_excl_scan4((nlo << 16) | nib, part);      // part[0..3]: free again
        uint32_t ib = pre & 0xFFFFu;
        const uint32_t sel0 = (pre >> 16) + min(ib, quota);
        uint32_t selm = 0;
#pragma unroll
        for (uint32_t u = 0; u < kSelPerThread; ++u) {
            if (i0 + u >= c) continue;
            const uint32_t k = ek[u];
            const bool go = k < lo_base || (k < thr_hi && ib++ < quota);
            if (go) selm |= 1u << u;
        }
        const uint32_t nsel = (uint32_t)__popc(selm);
        if (blockIdx.x < nsolve) {
            // Selected tile of rank blockIdx.x + m * nsolve (list order) is this group's job m.
            const uint32_t njobs = blockIdx.x < n ? (n - 1u - blockIdx.x) / nsolve + 1u : 0u;
            for (uint32_t m = 0; m < njobs; ++m) {
                const uint32_t r = blockIdx.x + m * nsolve;
                if (r >= sel0 && r < sel0 + nsel) {
                    const uint32_t want = r - sel0;
                    uint32_t got = 0u;
#pragma unroll
                    for (uint32_t u = 0; u < kSelPerThread; ++u) {
                        if ((selm >> u) & 1u) {
                            if (got == want) s_jobs[m] = eg[u];
                            ++got;
                        }
                    }
                }
            }
            __syncthreads();
            for (uint32_t m = 0; m < njobs; ++m) {
                const uint32_t g = s_jobs[m];
                if (g >= a.tiles) continue;            // never taken: every rank has one owner
                sweep_tile<TW, TH, KIND, false>(a, g, wave, nxt, np, blockIdx.x + m * nsolve, sh, s_u32, s_f32, in0);
                __syncthreads();
            }
        } else {
            // Carry group q takes the unselected entries of carry rank = q mod kSelCarryGroups.
            const uint32_t q = blockIdx.x - nsolve;
#pragma unroll
            for (uint32_t u = 0; u < kSelPerThread; ++u) {
                const uint32_t i = i0 + u;
                if (i >= c || ((selm >> u) & 1u)) continue;
                const uint32_t r = i - sel0 - (uint32_t)__popc(selm & ((1u << u) - 1u));
                if (r % kSelCarryGroups == q)
                    activate_np(a, eg[u], wave, nxt, nxt, kMaskOpenAll, __uint_as_float(ek[u]));
            }
        }
    }
}


// Wave boundary, run by one work-group of NT threads: empty the consumed list and swap
// (unless `first`), compute the shard prefix of the new list and, when it holds more than
// n tiles, select its n lowest keys: a two-level radix histogram over key bits [30:19]
// then [18:7] (keys are non-negative floats, so their bit patterns order like the values)
// gives the threshold bin; entries below it and, up to the quota, inside it are compacted
// into jobs[]; the rest are carried to the next list with their key and border bits.
// Returns the new current list, or -1 without writing anything when `stop_if_done` is set
// and the loop is already done.
template <int NT, int STAGE, bool HANDOFF = false>
__device__ int wave_select(const SolveArgs& a, bool first, uint32_t* hist, uint32_t* part,
                           uint32_t* s_prefix, uint32_t* s_misc, uint32_t* sg, uint32_t* sk,
                           bool stop_if_done = false) {
    static_assert(kShards == 32 && NT >= 64, "one wave64 holds the shard counts of both lists");
    GhcmCtrl* ctrl = a.ctrl;
    const int tid = threadIdx.x;
    // The control words and the shard counts of both lists in one round trip.
    const uint32_t done = stop_if_done ? ctrl->done : 0u;
    const uint32_t cur0 = ctrl->cur;
    const uint32_t wave0 = ctrl->wave;
    const uint32_t n = ctrl->n_select;
    // Read here, with the other words, and written back below without a second round trip:
    // nothing else writes it while a boundary runs.
    const uint32_t busy0 = ctrl->busy_waves;
    uint32_t cnt = 0u;
    if (tid < 2 * (int)kShards) cnt = ctrl->count[tid >> 5][(tid & 31) * kCounterStride];
    // Keeps the compiler from sinking these reads (and the `first` argument) below the done
    // test, which would cost a second round trip on every boundary.  The statement is empty.
    asm volatile("" ::"s"(busy0), "s"((uint32_t)first));
    if (done) return -1;
    // list[cur0] was consumed by the wave that just ran (unless `first`): it is emptied and
    // the other list becomes current.
    const uint32_t cur = first ? cur0 : (cur0 ^ 1u);
    if (tid < 64) {
        // Lanes 0-31 scan list 0's counts, lanes 32-63 list 1's (wave 0, every lane active).
        const uint32_t sl = (uint32_t)tid & 31u, lst = (uint32_t)tid >> 5;
        const uint32_t x = half_incl_scan_dpp(cnt);
        if (lst == cur) {
            s_prefix[sl + 1] = x;
            if (sl == 0) s_prefix[0] = 0u;
        }
        if (!first && lst == cur0) ctrl->count[cur0][sl * kCounterStride] = 0u;
    }
    __syncthreads();
    const uint32_t c = s_prefix[kShards];
    const uint32_t wave = wave0 + (first ? 0u : 1u);
    const uint32_t p = wave & 1u;
    if (tid <= (int)kShards) ctrl->prefix[tid] = s_prefix[tid];
    if (tid == 0) {
        ctrl->cur = cur;
        ctrl->wave = wave;
        ctrl->total = c;
        if (c == 0u) ctrl->done = 1u;
        else ctrl->busy_waves = busy0 + 1u;
        ctrl->jobs_from_list = 1u;
        ctrl->n_jobs = c;
        ctrl->n_carry = 0u;
    }
    if (c <= n) return (int)cur;

    const uint32_t* key = a.key + (size_t)p * a.tiles;
    // Entries (tile, key) are read from the list once into LDS when they fit.
    const bool staged = STAGE > 0 && c <= (uint32_t)STAGE;
    if (staged) {
        for (uint32_t i = tid; i < c; i += NT) {
            const uint32_t g = list_entry(a, s_prefix, cur, i);
            sg[i] = g;
            sk[i] = key[g];
        }
        __syncthreads();
    }
    // Two copies of the histogram and compaction passes, one reading the LDS stage and one
    // the list in HBM: a runtime choice between the two would make every read a flat one.
    uint32_t base = 0, carried = 0;
    bool handed = false;                 // carried tiles left to the tile launch (HANDOFF)
    auto passes = [&](auto staged_c) {
        constexpr bool S = decltype(staged_c)::value;
        auto entry_g = [&](uint32_t i) { return S ? sg[i] : list_entry(a, s_prefix, cur, i); };
        auto entry_k = [&](uint32_t i, uint32_t g) { return S ? sk[i] : key[g]; };
        uint32_t lo_base = 0u, below = 0u;
        constexpr int BPT = 4096 / NT;                    // bins per thread
        for (int level = 0; level < 2; ++level) {
            const int shift = level == 0 ? 19 : 7;
            for (int i = tid; i < 4096; i += NT) hist[i] = 0u;
            __syncthreads();
            for (uint32_t i = tid; i < c; i += NT) {
                const uint32_t k = entry_k(i, S ? 0u : list_entry(a, s_prefix, cur, i));
                if (level == 1 && (k >> 19) != (lo_base >> 19)) continue;
                atomicAdd(&hist[(k >> shift) & 0xFFFu], 1u);
            }
            __syncthreads();
            uint32_t mine = 0;
            for (int j = 0; j < BPT; ++j) mine += hist[BPT * tid + j];
            uint32_t tot;
            const uint32_t excl = block_excl_scan<NT>(mine, part, &tot) + below;
            if (excl < n && excl + mine >= n) {           // the crossing bin is one of mine
                uint32_t cum = excl, b = BPT * tid;
                for (int j = 0; j < BPT; ++j) {
                    const uint32_t h = hist[BPT * tid + j];
                    if (cum + h >= n) { b = BPT * tid + j; break; }
                    cum += h;
                }
                s_misc[0] = b;
                s_misc[1] = cum;
            }
            __syncthreads();
            lo_base |= s_misc[0] << shift;
            below = s_misc[1];
            __syncthreads();
        }
        // Compaction: keys < lo_base (there are `below` of them) all go; keys in
        // [lo_base, lo_base + 128) go, in list order, while the quota n - below lasts.  Jobs
        // keep list order.  The histogram is free from here on; its first kShards words count
        // the carried entries per shard.
        if (tid < (int)kShards) hist[tid] = 0u;
        __syncthreads();
        const uint32_t quota = n - below;
        const uint32_t thr_hi = lo_base + (1u << 7);
        const size_t np = p ^ 1u;
        // Carry: move the activation record to the next parity and queue the tile in list
        // cur^1 for the wave after.  That list was emptied above (or is still empty on the
        // first boundary) and nothing else runs during the boundary, so the tile is not in it
        // (qwave < wave + 2), its shard slot comes from an LDS counter and qwave is a plain
        // store: no returned global atomic sits on this path (activate() needs two in a row).
        auto carry = [&](uint32_t g, uint32_t k, uint32_t mb) {
            a.key[(size_t)p * a.tiles + g] = kInfBits;
            a.mask[(size_t)p * a.tiles + g] = 0u;
            atomicMin(&a.key[np * a.tiles + g], k);
            atomicOr(&a.mask[np * a.tiles + g], mb);
            a.qwave[g] = wave + 2u;
            const uint32_t sh = g % kShards;
            const uint32_t slot = atomicAdd(&hist[sh], 1u);
            a.list[((size_t)(cur ^ 1u) * kShards + sh) * a.shard_cap + slot] = g;
        };
        if (S && c > 2u * NT) {
            // Long staged lists: each thread takes a contiguous run of `per` entries, so two
            // block scans (ties, then selections) cover the whole list.  With HANDOFF (after
            // a solve's first boundary) the carried tiles are written after the selected ones
            // for the tile launch to move
            // (its whole grid does that in parallel); otherwise their border bits are fetched
            // as one batch of loads and this work-group moves them.
            // (Lists of up to two block passes keep the loop below, which times better there.)
            constexpr int UM = STAGE > 0 ? (STAGE + NT - 1) / NT : 1;
            const uint32_t per = (c + NT - 1) / NT;       // <= UM since c <= STAGE
            const uint32_t i0 = (uint32_t)tid * per;
            uint32_t ntie = 0;
            for (uint32_t u = 0; u < per; ++u) {
                const uint32_t i = i0 + u;
                if (i < c) {
                    const uint32_t k = sk[i];
                    ntie += (k >= lo_base && k < thr_hi) ? 1u : 0u;
                }
            }
            uint32_t tot_tie;
            uint32_t rank = block_excl_scan<NT>(ntie, part, &tot_tie);
            uint32_t selbits = 0, nsel = 0;
            for (uint32_t u = 0; u < per; ++u) {
                const uint32_t i = i0 + u;
                if (i < c) {
                    const uint32_t k = sk[i];
                    const bool low = k < lo_base;
                    const bool tie = !low && k < thr_hi;
                    if (low || (tie && rank < quota)) {
                        selbits |= 1u << u;
                        ++nsel;
                    }
                    rank += tie ? 1u : 0u;
                }
            }
            uint32_t tot_sel;
            uint32_t pos = block_excl_scan<NT>(nsel, part, &tot_sel);
            // Not on a solve's first boundary: the persistent and fused forms start with one,
            // and their tile loops do not take handed-off tiles.
            if (HANDOFF && !first) {
                // Entries before this thread's run that were not selected come first.
                uint32_t cpos = tot_sel + min(i0, c) - pos;
                for (uint32_t u = 0; u < per; ++u) {
                    const uint32_t i = i0 + u;
                    if (i < c) {
                        if ((selbits >> u) & 1u) a.jobs[pos++] = sg[i];
                        else a.jobs[cpos++] = sg[i];
                    }
                }
                handed = true;
            } else {
                uint32_t mbv[UM];
#pragma unroll
                for (int u = 0; u < UM; ++u) {
                    const uint32_t i = i0 + (uint32_t)u;
                    mbv[u] = 0u;
                    if ((uint32_t)u < per && i < c && !((selbits >> u) & 1u))
                        mbv[u] = a.mask[(size_t)p * a.tiles + sg[i]];
                }
#pragma unroll
                for (int u = 0; u < UM; ++u) {
                    const uint32_t i = i0 + (uint32_t)u;
                    if ((uint32_t)u < per && i < c) {
                        const uint32_t g = sg[i];
                        if ((selbits >> u) & 1u) a.jobs[pos++] = g;
                        else carry(g, sk[i], mbv[u]);
                    }
                }
            }
            base = tot_sel;
            carried = c - tot_sel;
        } else {
            uint32_t quota_used = 0;
            for (uint32_t i0 = 0; i0 < c; i0 += NT) {
                const uint32_t i = i0 + tid;
                uint32_t g = 0, k = 0xFFFFFFFFu, low = 0, tie = 0;
                if (i < c) {
                    g = entry_g(i);
                    k = entry_k(i, g);
                    low = k < lo_base ? 1u : 0u;
                    tie = (!low && k < thr_hi) ? 1u : 0u;
                }
                uint32_t tot_tie;
                const uint32_t tie_rank = block_excl_scan<NT>(tie, part, &tot_tie) + quota_used;
                const uint32_t sel = low | ((tie && tie_rank < quota) ? 1u : 0u);
                uint32_t tot_sel;
                const uint32_t pos = block_excl_scan<NT>(sel, part, &tot_sel);
                if (i < c) {
                    if (sel) a.jobs[base + pos] = g;
                    else carry(g, k, a.mask[(size_t)p * a.tiles + g]);
                }
                base += tot_sel;
                quota_used += tot_tie;
                carried += (i0 + NT <= c ? NT : c - i0) - tot_sel;
            }
        }
    };
    if (staged) passes(std::true_type{});
    else passes(std::false_type{});
    __syncthreads();
    if (tid < (int)kShards && hist[tid]) ctrl->count[cur ^ 1u][tid * kCounterStride] = hist[tid];
    if (tid == 0) {
        ctrl->jobs_from_list = 0u;
        ctrl->n_jobs = base;
        ctrl->carried += carried;
        if (handed) ctrl->n_carry = carried;
    }
    return (int)cur;
}

constexpr int kSelectStage = 16384;    // list entries staged in LDS by k_wave_end (128 KB)

// Serving: queued tiles per instance slot in the new current list `cur` (prefix in
// s_prefix), tagged with the slot's admission generation; a count of 0 means nothing is
// left to solve.  `cnt` is an LDS array of at least n_slots words.
template <int NT>
__device__ void publish_slots(const SolveArgs& a, const uint32_t* s_prefix, uint32_t cur,
                              uint32_t* cnt) {
    const int tid = threadIdx.x;
    const uint32_t ns = a.n_slots;
    for (uint32_t i = tid; i < ns; i += NT) cnt[i] = 0u;
    __syncthreads();
    const uint32_t c = s_prefix[kShards];
    for (uint32_t i = tid; i < c; i += NT)
        atomicAdd(&cnt[a.tinst[list_entry(a, s_prefix, cur, i)]], 1u);
    __syncthreads();
    for (uint32_t i = tid; i < ns; i += NT)
        sys_store64(&a.rec->slot[i], ((unsigned long long)a.sgen[i] << 32) | cnt[i]);
}

__global__ __launch_bounds__(1024) void k_wave_end(SolveArgs a, uint32_t first) {
    __shared__ uint32_t hist[4096];
    __shared__ uint32_t part[128];
    __shared__ uint32_t s_prefix[kShards + 1];
    __shared__ uint32_t s_misc[2];
    __shared__ uint32_t sg[kSelectStage];
    __shared__ uint32_t sk[kSelectStage];
    const int tid = threadIdx.x;
    // Read together with the control words (one round trip); only this kernel writes it.
    const uint32_t ticks0 = a.ctrl->ticks;
    const int cur = wave_select<1024, kSelectStage, true>(a, first != 0u, hist, part, s_prefix,
                                                          s_misc, sg, sk, true);
    // An idle launch (the loop is done) only advances the tick, so the host still sees progress.
    if (tid == 0) a.ctrl->ticks = ticks0 + 1u;
    if (cur < 0 || !a.rec || !a.n_slots) return;
    __syncthreads();
    publish_slots<1024>(a, s_prefix, (uint32_t)cur, hist);
}

// Per-solve initialisation in one launch (grid.y over instances): the control block and
// totals zeroed, the instance table copied from mapped host memory, every tile's queue
// state reset and mapped to its instance, every time field set to +INF.
// Edge columns of one instance's tiles (the sweep kernels' halo source; eh = th * td rows per
// column, row r at (y, z) = (r mod th, r / th) of the tile): time +INF and the cost of the same
// cells (+INF past the grid).  Entry e of the instance is row e mod eh of tile e / (2 eh)'s left
// column, or of its right one when e mod 2eh >= eh.  Threads [first, ...) with `stride`.
template <int KIND>
__device__ void init_edges(const SolveArgs& a, const GhcmInst& in, uint32_t nt, uint32_t tw,
                           uint32_t th, uint32_t td, uint32_t first, uint32_t stride) {
    const uint32_t eh = th * td, e2 = 2u * eh;
    const uint32_t ne = nt * e2;
    for (uint32_t e = first; e < ne; e += stride) {
        const uint32_t t = e / e2, j = e - t * e2;
        const uint32_t right = j >= eh ? 1u : 0u, r = j - right * eh;
        const uint32_t tx = t % in.tiles_x, ty = (t / in.tiles_x) % in.tiles_y;
        const uint32_t tz = t / (in.tiles_x * in.tiles_y);
        const uint32_t x = tx * tw + (right ? tw - 1 : 0u), y = ty * th + r % th, z = tz * td + r / th;
        float c = __builtin_inff();
        if (x < in.w && y < in.h && z < in.d)
            c = load_cost_at<KIND>(static_cast<const char*>(in.field) +
                                   (((size_t)z * in.h + y) * in.w + x) * field_bytes<KIND>());
        const size_t o = (size_t)(in.tile_base + t) * e2 + j;
        gmem(a.edge)[o] = __builtin_inff();
        gmem(const_cast<float*>(a.cedge))[o] = c;
    }
}

// tw != 0 (with a.edge set): the edge columns of every tile are initialised too, each
// instance's by the work-groups that fill its time field (no instance lookup per tile).
template <int KIND>
__global__ void k_solve_init(SolveArgs a, const GhcmInst* h_inst, uint32_t n_inst,
                             uint32_t n_select, unsigned long long* totals, uint32_t tw,
                             uint32_t th, uint32_t td) {
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        uint32_t* c = reinterpret_cast<uint32_t*>(a.ctrl);
        for (uint32_t i = threadIdx.x; i < sizeof(GhcmCtrl) / 4; i += blockDim.x) c[i] = 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            a.ctrl->n_select = n_select;
            totals[0] = totals[1] = totals[2] = 0ull;
        }
    }
    const size_t T = a.tiles;
    for (uint32_t j = blockIdx.y; j < n_inst; j += gridDim.y) {
        const GhcmInst in = h_inst[j];
        if (blockIdx.x == 0 && threadIdx.x == 0) const_cast<GhcmInst*>(a.inst)[j] = in;
        const uint32_t nt = inst_tiles(in);
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
            const size_t g = in.tile_base + t;
            a.qwave[g] = 0u;
            a.key[g] = kInfBits;
            a.key[T + g] = kInfBits;
            a.mask[g] = 0u;
            a.mask[T + g] = 0u;
            a.tmin[g] = __builtin_inff();
            a.tstat[g] = 0u;
            a.tstat[T + g] = 0u;
            a.tstat[2 * T + g] = 0u;
            if (a.tinst) const_cast<uint32_t*>(a.tinst)[g] = j;
        }
        auto* tm = gmem(reinterpret_cast<uint32_t*>(in.time));
        const size_t cells = (size_t)in.w * in.h * in.d;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += stride)
            tm[i] = kInfBits;
        if (tw && a.edge)
            init_edges<KIND>(a, in, nt, tw, th, td, blockIdx.x * blockDim.x + threadIdx.x, stride);
    }
}

// Totals of the per-tile counters handed to the host record; the tag is written last.
__global__ void k_publish_totals(const unsigned long long* totals, const GhcmCtrl* ctrl,
                                 HostRec* rec, uint32_t tag) {
    if (threadIdx.x == 0) {
        for (int j = 0; j < 3; ++j) sys_store64(&rec->totals[j], totals[j]);
        sys_store(&rec->wave, ctrl->wave);
        sys_store(&rec->busy_waves, ctrl->busy_waves);
        sys_store(&rec->totals_tag, tag);
    }
}

// Bulk copy between device memory and mapped host memory (or device to device), 16 bytes
// per lane where both ends are 16-byte aligned.  Used in place of runtime copies so no
// copy engine sits between the host and the wave loop.
__global__ void k_copy(const void* src, void* dst, size_t bytes) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const bool vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0;
    size_t done = 0;
    if (vec) {
        const size_t n16 = bytes / 16;
        const uint4* s4 = static_cast<const uint4*>(src);
        uint4* d4 = static_cast<uint4*>(dst);
        for (size_t i = tid; i < n16; i += stride) d4[i] = s4[i];
        done = n16 * 16;
    }
    const unsigned char* s1 = static_cast<const unsigned char*>(src);
    unsigned char* d1 = static_cast<unsigned char*>(dst);
    for (size_t i = done + tid; i < bytes; i += stride) d1[i] = s1[i];
}

// Tile solve and wave boundary in one launch (launch_mode 3): every work-group solves its
// jobs as k_solve_sweep does, makes its writes visible device-wide (agent-scope fence)
// and arrives on a counter; the last to arrive runs the wave boundary of k_wave_end --
// list swap, top-n selection, tick -- with its 256 threads and the tile's LDS.  One
// launch per wave instead of two.  The list entries the selection reads are appended by
// wave 0 of each work-group (threads 0-3, and thread 0 for a capped tile), so thread 0's
// fence orders them before the arrival.
constexpr int kFusedStage = 2048;
template <int TW, int TH>
constexpr int fused_lds_words() {
    return sweep_lds_words<TW, TH>() > 4096 + 128 + 2 * kFusedStage ? sweep_lds_words<TW, TH>()
                                                                    : 4096 + 128 + 2 * kFusedStage;
}

template <int TW, int TH, int KIND>
__global__ __launch_bounds__(kSweepThreads) void k_solve_sweep_fused(SolveArgs a) {
    __shared__ float sh[fused_lds_words<TW, TH>()];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[8];
    __shared__ uint32_t s_prefix[kShards + 1];
    __shared__ uint32_t s_misc[4];
    const int tid = threadIdx.x;
    const GhcmInst in0 = single_inst(a);
    GhcmCtrl* ctrl = a.ctrl;
    publish_progress(a, ctrl);
    const uint32_t done = ctrl->done;
    const uint32_t nj = ctrl->n_jobs;
    const uint32_t wave = ctrl->wave;
    const uint32_t cur = ctrl->cur;
    const uint32_t from_list = ctrl->jobs_from_list;
    // Tiles a wave boundary handed to this launch (none unless a k_wave_end after a solve's
    // first boundary selected from a long list) are carried like in k_solve_sweep.
    const uint32_t ncarry = ctrl->n_carry;
    if (!done && (blockIdx.x < nj || ncarry != 0u)) {
        const uint32_t p = wave & 1u;
        for (int i = tid; i <= (int)kShards; i += kSweepThreads) s_prefix[i] = ctrl->prefix[i];
        __syncthreads();
        for (uint32_t j = blockIdx.x; j < nj; j += gridDim.x) {
            const uint32_t g = job_tile(a, s_prefix, cur, from_list, j);
            if (tid == 0) claim_tile(a, g, p);
            sweep_tile<TW, TH, KIND, false>(a, g, wave, cur ^ 1u, p ^ 1u, j, sh, s_u32, s_f32, in0);
            __syncthreads();
        }
        if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
    }
    __syncthreads();
    if (tid == 0) {
        // Release only (L2 write-back, no invalidate): the other work-groups on this XCD
        // keep their cached lines; only the last arrival acquires.
        const uint32_t prev = __hip_atomic_fetch_add(&ctrl->arrive_top, 1u, __ATOMIC_RELEASE,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        s_misc[3] = (prev == gridDim.x - 1u) ? 1u : 0u;
    }
    __syncthreads();
    if (!s_misc[3]) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (tid == 0) ctrl->arrive_top = 0u;
    if (done) {
        if (tid == 0) ctrl->ticks += 1u;
        return;
    }
    uint32_t* hist = reinterpret_cast<uint32_t*>(sh);
    uint32_t* part = hist + 4096;
    uint32_t* sg = part + 128;
    uint32_t* sk = sg + kFusedStage;
    wave_select<kSweepThreads, kFusedStage>(a, false, hist, part, s_prefix, s_misc, sg, sk);
    if (tid == 0) ctrl->ticks += 1u;
    if (!a.rec || !a.n_slots) return;
    __syncthreads();
    publish_slots<kSweepThreads>(a, s_prefix, cur ^ 1u, hist);
}

// A seed on a tile's first or last column also sets that tile's edge column.
__device__ __forceinline__ void seed_edges(const SolveArgs& a, const GhcmInst& in, uint32_t x,
                                           uint32_t y, uint32_t tw, uint32_t th) {
    if (!a.edge) return;
    const uint32_t t = tile_of_cell(in.tile_base, in.tiles_x, x, y, tw, th);
    const uint32_t cx = x % tw, r = y % th;
    if (cx == 0) gmem(a.edge)[(size_t)t * 2 * th + r] = 0.0f;
    if (cx == tw - 1) gmem(a.edge)[(size_t)t * 2 * th + th + r] = 0.0f;
}

// Serving admission: seeds of newly admitted instances, (x, y, slot) triples.  Queues the
// tiles of each seed and its 4 neighbours for the wave after the current one (they join
// list[cur^1] like any activation) and clears `done` so an idle loop resumes.
__global__ void k_seed_serve(SolveArgs a, const uint32_t* seeds, uint32_t n_seeds,
                             uint32_t tw, uint32_t th) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seeds) return;
    const uint32_t wave = a.ctrl->wave;
    const uint32_t cur = a.ctrl->cur;
    const uint32_t np = (wave + 1u) & 1u;
    const uint32_t x = seeds[3 * s], y = seeds[3 * s + 1], ii = seeds[3 * s + 2];
    const GhcmInst in = a.inst[ii];
    gmem(in.time)[(size_t)y * in.w + x] = 0.0f;
    seed_edges(a, in, x, y, tw, th);
    const int dx[5] = {0, -1, 1, 0, 0}, dy[5] = {0, 0, 0, -1, 1};
    for (int d = 0; d < 5; ++d) {
        const int px = (int)x + dx[d], py = (int)y + dy[d];
        if (px < 0 || py < 0 || px >= (int)in.w || py >= (int)in.h) continue;
        const uint32_t t = tile_of_cell(in.tile_base, in.tiles_x, (uint32_t)px, (uint32_t)py, tw, th);
        atomicMin(&a.key[(size_t)np * a.tiles + t], 0u);
        atomicOr(&a.mask[(size_t)np * a.tiles + t], kMaskOpenAll);
        const uint32_t old = atomicMax(&a.qwave[t], wave + 2u);
        if (old < wave + 2u) list_append(a, cur ^ 1u, t);
    }
}

// Serving admission, last step: the loop leaves its idle state so the next wave boundary
// reports the new slots (with or without seeds).
__global__ void k_wake(GhcmCtrl* ctrl) {
    if (threadIdx.x == 0) ctrl->done = 0u;
}

// Serving admission in one launch: the slot's instance entry written, its tile range reset
// and mapped to the slot, its time field set to +INF and (tw != 0, sweep kernel) its tiles'
// edge columns initialised from the instance record passed by value.
template <int KIND>
__global__ void k_admit_tiles(SolveArgs a, GhcmInst in, uint32_t slot, uint32_t gen, uint32_t tw,
                              uint32_t th) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const_cast<GhcmInst*>(a.inst)[slot] = in;
        const_cast<uint32_t*>(a.sgen)[slot] = gen;
    }
    const uint32_t base = in.tile_base, count = inst_tiles(in);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t t = base + i;
        a.qwave[t] = 0u;
        a.key[t] = kInfBits;
        a.key[(size_t)a.tiles + t] = kInfBits;
        a.mask[t] = 0u;
        a.mask[(size_t)a.tiles + t] = 0u;
        a.tmin[t] = __builtin_inff();
        a.tstat[t] = 0u;
        a.tstat[(size_t)a.tiles + t] = 0u;
        a.tstat[2 * (size_t)a.tiles + t] = 0u;
        const_cast<uint32_t*>(a.tinst)[t] = slot;
    }
    // The request's time field at +INF and, for the sweep kernel, its tiles' edge columns.
    const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    auto* tm = gmem(reinterpret_cast<uint32_t*>(in.time));
    const size_t cells = (size_t)in.w * in.h * in.d;
    for (size_t i = first; i < cells; i += stride) tm[i] = kInfBits;
    if (tw && a.edge) init_edges<KIND>(a, in, count, tw, th, 1u, first, stride);
}

// ---------------------------------------------------------------------------------------
// Persistent device wave loop for the sweep solver: one cooperative launch runs every
// wave.  Each work-group solves its share of the wave's jobs, then arrives at a
// two-level counter (groups of kBarGroup work-groups, then a top counter); the last
// arrival runs wave_select for the next wave with its whole work-group and publishes the
// new wave index; the others wait for it.  Releases are agent-scope fences by every
// thread before arriving; the waiter's acquire fence precedes its reads of the new wave.
// A wait longer than ~2^26 polls sets `abort` and every work-group leaves the loop.
// ---------------------------------------------------------------------------------------
template <int TW, int TH, int KIND>
__global__ __launch_bounds__(kSweepThreads) void k_persistent_sweep(SolveArgs a) {
    constexpr int NW = sweep_lds_words<TW, TH>();
    constexpr int TS = NW > 4096 + kSweepThreads ? NW : 4096 + kSweepThreads;
    // The selection's histogram and scan buffer reuse the tile buffer: the work-group
    // that runs wave_select has finished its tiles for the wave.
    __shared__ float sh[TS];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[8];
    __shared__ uint32_t s_prefix[kShards + 1];
    __shared__ uint32_t s_misc[4];
    uint32_t* hist = reinterpret_cast<uint32_t*>(sh);
    uint32_t* part = reinterpret_cast<uint32_t*>(sh) + 4096;
    GhcmCtrl* ctrl = a.ctrl;
    const GhcmInst in0 = single_inst(a);
    const int tid = threadIdx.x;
    const uint32_t G = gridDim.x;
    const uint32_t n_groups = (G + kBarGroup - 1) / kBarGroup;
    const uint32_t grp = blockIdx.x / kBarGroup;
    const uint32_t grp_size = min(kBarGroup, G - grp * kBarGroup);
    for (;;) {
        if (tid == 0) {
            s_misc[2] = __hip_atomic_load(&ctrl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
                        __hip_atomic_load(&ctrl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (s_misc[2]) break;
        const uint32_t wave = ctrl->wave;
        const uint32_t cur = ctrl->cur;
        const uint32_t nj = ctrl->n_jobs;
        const uint32_t from_list = ctrl->jobs_from_list;
        const uint32_t ncarry = ctrl->n_carry;
        const uint32_t p = wave & 1u;
        for (int i = tid; i <= (int)kShards; i += kSweepThreads) s_prefix[i] = ctrl->prefix[i];
        __syncthreads();
        for (uint32_t j = blockIdx.x; j < nj; j += G) {
            const uint32_t g = job_tile(a, s_prefix, cur, from_list, j);
            if (tid == 0) claim_tile(a, g, p);
            sweep_tile<TW, TH, KIND, false>(a, g, wave, cur ^ 1u, p ^ 1u, j, sh, s_u32, s_f32, in0);
            __syncthreads();
        }
        // Tiles a boundary handed to the tile loop (none from this kernel's own selections,
        // which carry directly) are carried by the whole grid.
        if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
        // ---- arrive ----
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            uint32_t last = 0;
            const uint32_t o = atomicAdd(&ctrl->arrive[grp * kCounterStride], 1u);
            if (o == grp_size - 1u) {
                const uint32_t t = atomicAdd(&ctrl->arrive_top, 1u);
                last = (t == n_groups - 1u) ? 1u : 0u;
            }
            s_misc[3] = last;
        }
        __syncthreads();
        if (s_misc[3]) {
            __threadfence();
            wave_select<kSweepThreads, 0>(a, false, hist, part, s_prefix, s_misc, nullptr, nullptr);
            if (tid < (int)n_groups) ctrl->arrive[tid * kCounterStride] = 0u;
            if (tid == 0) ctrl->arrive_top = 0u;
            __threadfence();
            __syncthreads();
            if (tid == 0)
                __hip_atomic_store(&ctrl->wave_ready, wave + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else if (tid == 0) {
            uint32_t polls = 0;
            while (__hip_atomic_load(&ctrl->wave_ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != wave + 1u) {
                if (__hip_atomic_load(&ctrl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (++polls > (1u << 26)) {
                    atomicOr(&ctrl->abort, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
}

// Sums the per-tile counters into out[0..2] (visits, passes, improved cells).
__global__ void k_reduce_stats(const uint32_t* tstat, uint32_t tiles, unsigned long long* out) {
    __shared__ unsigned long long red[3][256];
    unsigned long long v[3] = {0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tiles;
         i += (size_t)gridDim.x * blockDim.x) {
        v[0] += tstat[i];
        v[1] += tstat[(size_t)tiles + i];
        v[2] += tstat[2 * (size_t)tiles + i];
    }
    for (int j = 0; j < 3; ++j) red[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 3; ++j) atomicAdd(&out[j], red[j][0]);
}

// ---- initialisation kernels (fill kernels rather than memsets) ----
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_tile_inst(uint32_t* tinst, const GhcmInst* inst, uint32_t n_inst) {
    for (uint32_t j = blockIdx.x; j < n_inst; j += gridDim.x) {
        const GhcmInst in = inst[j];
        const uint32_t nt = inst_tiles(in);
        for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) tinst[in.tile_base + t] = j;
    }
}

// Seeds: T = 0 at each seed, and the tiles of the seed and its 4 neighbours are
// queued for wave 0 with every cell open (src/eikonal_hcm.cpp:189-238).
// seeds: (x, y, instance) triples.
__global__ void k_seed(SolveArgs a, const uint32_t* seeds, uint32_t n_seeds,
                       uint32_t tw, uint32_t th) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seeds) return;
    const uint32_t x = seeds[3 * s], y = seeds[3 * s + 1], ii = seeds[3 * s + 2];
    const GhcmInst in = a.inst[ii];
    gmem(in.time)[(size_t)y * in.w + x] = 0.0f;
    seed_edges(a, in, x, y, tw, th);
    const int dx[5] = {0, -1, 1, 0, 0}, dy[5] = {0, 0, 0, -1, 1};
    for (int d = 0; d < 5; ++d) {
        int px = (int)x + dx[d], py = (int)y + dy[d];
        if (px < 0 || py < 0 || px >= (int)in.w || py >= (int)in.h) continue;
        const uint32_t t = tile_of_cell(in.tile_base, in.tiles_x, (uint32_t)px, (uint32_t)py, tw, th);
        atomicMin(&a.key[t], 0u);
        atomicOr(&a.mask[t], kMaskOpenAll);
        uint32_t old = atomicMax(&a.qwave[t], 1u);
        if (old < 1u) list_append(a, 0u, t);
    }
}

// ---- image processing: RGB -> luminosity (data/img_proc_rgb_luminosity.cl:5-19) ----
// out = dot(rgb, (0.299, 0.587, 0.114)) per pixel, evaluated as
// fma(b, 0.114, fma(g, 0.587, r * 0.299)).  HBM-bound: 12 bytes in, 4 out per pixel.
// Each thread takes four pixels as three 16-byte loads and one 16-byte store when the
// buffers are 16-byte aligned; the tail (and unaligned buffers) go pixel by pixel.
__device__ __forceinline__ float luminosity(float r, float g, float b) {
    return fmaf(b, 0.114f, fmaf(g, 0.587f, r * 0.299f));
}

__global__ void k_rgb_luminosity(const float* __restrict__ rgb, float* __restrict__ out, size_t n) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const bool vec = ((reinterpret_cast<uintptr_t>(rgb) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    size_t done = 0;
    if (vec) {
        const size_t n4 = n / 4;
        const float4* in4 = reinterpret_cast<const float4*>(rgb);
        float4* out4 = reinterpret_cast<float4*>(out);
        for (size_t i = tid; i < n4; i += stride) {
            const float4 a = in4[3 * i], b = in4[3 * i + 1], c = in4[3 * i + 2];
            // a = r0 g0 b0 r1, b = g1 b1 r2 g2, c = b2 r3 g3 b3
            out4[i] = make_float4(luminosity(a.x, a.y, a.z), luminosity(a.w, b.x, b.y),
                                  luminosity(b.z, b.w, c.x), luminosity(c.y, c.z, c.w));
        }
        done = n4 * 4;
    }
    for (size_t i = done + tid; i < n; i += stride)
        out[i] = luminosity(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
}

// ---------------------------------------------------------------------------------------
// Three-dimensional solve (K = 3): W x H x D grids in 8 x 8 x 8 tiles
// ---------------------------------------------------------------------------------------

// Godunov upwind update in three dimensions for |grad T| = f, h = 1 (the K-axis update of
// src/eikonal.h:73-118 and data/eikonal_common.cl:221-285 at K = 3), on the three axis
// minima a, b, c in any order.  Sorted m1 <= m2 <= m3, the update is the first of
//   m1 + f                                    if it does not exceed m2,
//   the two-axis root                         if it does not exceed m3,
//   m1 + (u + v + sqrt(3 f^2 - u^2 - v^2 - (u - v)^2)) / 3, u = m2 - m1, v = m3 - m1,
// the larger root of (T-m1)^2 + (T-m2)^2 + (T-m3)^2 = f^2 with the common part m1 taken out
// of the discriminant (the reference forms it from the unshifted values).
// Evaluated without a compare or select, by clamping each added axis to where the form
// below it stops being valid (the forms meet there):
//   t2 = the two-axis form at e = min(u, f) (as godunov_raw: exactly m1 + f for u >= f),
//   m3c = min(m3, t2): for m3 >= t2 the three-axis root at v = t2 - m1 is t2 itself,
//   t3 = the three-axis form at (e, m3c - m1),
//   T = med3(t3, m3c, t2), i.e. t3 clamped to [m3c, t2].
// For m3 >= t2 that is t2 exactly; for m3 < t2 it is t3, never below m3 (the largest minimum
// it used) and never above t2.  A float32 emulation over 3 M inputs spread over the three
// branches and their boundaries gives the branch form's result bit for bit except 32
// three-axis cases (1-2 ulp, where t3 had rounded above t2), and the same count of one-ulp
// non-monotone steps.  As in two dimensions, m1 is added by one final rounding.  +INF minima
// fall through to the lower-axis forms (an INF - INF difference makes a NaN that the clamps
// and the final med3, which returns the minimum of its other operands for a NaN operand, drop);
// f = +INF yields +INF.  Every cell update and every face prediction of both 3-D kernels goes
// through this one function.
__device__ __forceinline__ float godunov3(float a, float b, float c, float f) {
    // Sorted by three-operand minimum, median and maximum (v_min3/v_med3/v_max3: the values
    // are non-negative or +INF, never NaN, so the float forms order them exactly).
    const float m1 = fminf(fminf(a, b), c);
    const float m3 = fmaxf(fmaxf(a, b), c);
    const float m2 = __builtin_amdgcn_fmed3f(a, b, c);
    const float u = m2 - m1;
    const float s2 = 1.41421356237309515f * f;
    const float e = __builtin_amdgcn_fmed3f(u, f, 0.0f);
    const float r2 = __builtin_amdgcn_sqrtf((s2 - e) * (s2 + e));
    const float t2 = __builtin_fmaf(e + __builtin_fmaxf(r2, e), 0.5f, m1);
    const float m3c = __builtin_fminf(m3, t2);
    const float v = m3c - m1, w = e - v;
    const float disc = fmaxf(3.0f * f * f - (e * e + v * v + w * w), 0.0f);
    const float t3 = m1 + (e + v + __builtin_amdgcn_sqrtf(disc)) * (1.0f / 3.0f);
    return __builtin_amdgcn_fmed3f(t3, m3c, t2);
}

constexpr int kT3 = 8;                 // 3-D tile edge
constexpr int kT3Threads = 256;        // two cells per thread
constexpr int kT3Box = kT3 + 2;        // staged box edge (tile plus a one-cell halo)
constexpr uint32_t kT3HardPasses = 4 * kT3 * kT3 * kT3 + 64;
constexpr int kT3PassesPerVote = 2;

// Side numbering in three dimensions: 0 x-, 1 y-, 2 x+, 3 y+ (as in two), 4 z-, 5 z+.
__device__ __forceinline__ void side_step3(int s, int& dx, int& dy, int& dz) {
    dx = s == 0 ? -1 : s == 2 ? 1 : 0;
    dy = s == 1 ? -1 : s == 3 ? 1 : 0;
    dz = s == 4 ? -1 : s == 5 ? 1 : 0;
}

// One visit of 3-D tile g: stage the tile and its halo box in LDS, relax every cell in
// Jacobi passes (each thread owns two cells and updates them in place) until no cell
// improves by >= eps or the pass cap, predict the six neighbour faces, activate the
// neighbours whose face would improve, commit the improved cells.
template <int KIND>
__device__ void solve3d_tile(const SolveArgs& a, uint32_t g, uint32_t wave, uint32_t nxt,
                             uint32_t p, float* t_s, float* c_s, uint32_t* s_u32, float* s_f32,
                             GhcmInst in_single) {
    constexpr int B = kT3, PB = kT3Box, PB2 = PB * PB, PP = PB2 * PB, P = B * B * B;
    constexpr int CPT = P / kT3Threads;
    const int tid = threadIdx.x;
    GhcmInst in = in_single;
    if (a.n_inst != 1) in = uniform_inst(a.inst[a.tinst[g]]);
    const uint32_t lt = g - in.tile_base;
    const int tx = (int)(lt % in.tiles_x);
    const int ty = (int)((lt / in.tiles_x) % in.tiles_y);
    const int tz = (int)(lt / (in.tiles_x * in.tiles_y));
    const int W = (int)in.w, H = (int)in.h, D = (int)in.d;
    const int ox = tx * B, oy = ty * B, oz = tz * B;

    // Stage the PB^3 box (origin (ox-1, oy-1, oz-1)); out-of-grid cells read +INF.
    {
        constexpr int NL = (PP + kT3Threads - 1) / kT3Threads;
        float tv[NL], cv[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int i = tid + kT3Threads * k;
            const int bx = i % PB, by = (i / PB) % PB, bz = i / PB2;
            const int gx = ox + bx - 1, gy = oy + by - 1, gz = oz + bz - 1;
            tv[k] = __builtin_inff();
            cv[k] = __builtin_inff();
            if (i < PP && gx >= 0 && gy >= 0 && gz >= 0 && gx < W && gy < H && gz < D) {
                const size_t idx = ((size_t)gz * (size_t)H + (size_t)gy) * (size_t)W + (size_t)gx;
                tv[k] = gmem(in.time)[idx];
                cv[k] = load_cost_at<KIND>(static_cast<const char*>(in.field) + idx * field_bytes<KIND>());
            }
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int i = tid + kT3Threads * k;
            if (i < PP) {
                t_s[i] = tv[k];
                c_s[i] = cv[k];
            }
        }
    }
    __syncthreads();
    int q[CPT];
    float told[CPT], tcur[CPT], cost[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int i = tid + kT3Threads * k;
        const int x = i % B, y = (i / B) % B, z = i / (B * B);
        q[k] = ((z + 1) * PB + (y + 1)) * PB + (x + 1);
        told[k] = t_s[q[k]];
        tcur[k] = told[k];
        cost[k] = c_s[q[k]];
    }

    // In-tile passes.  Out-of-grid cells hold +INF time and cost and stay +INF.
    const float eps = a.eps;
    const uint32_t cap = (a.iter_cap == 0u || a.iter_cap > kT3HardPasses) ? kT3HardPasses : a.iter_cap;
    uint32_t passes = 0;
    bool unfinished = false;
    // Passes need no barrier between them (updates only lower cells, so a racing read sees an
    // upper bound); the block-wide convergence vote is taken every kT3PassesPerVote passes,
    // which cut the 256^3 volume solve from 5.08 to 4.82 ms (every 4: 4.87 ms).  A thread is
    // the only writer of its own cells during the passes, so their times (and costs) stay in
    // registers and only the six neighbours are read from LDS.
    for (;;) {
        bool changed = false;
#pragma unroll
        for (int rep = 0; rep < kT3PassesPerVote; ++rep) {
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                const int c = q[k];
                const float t0 = tcur[k];
                const float ax = fminf(t_s[c - 1], t_s[c + 1]);
                const float ay = fminf(t_s[c - PB], t_s[c + PB]);
                const float az = fminf(t_s[c - PB2], t_s[c + PB2]);
                const float tn = godunov3(ax, ay, az, cost[k]);
                if (tn < t0) {
                    t_s[c] = tn;
                    tcur[k] = tn;
                    if (t0 - tn >= eps) changed = true;
                }
            }
            ++passes;
        }
        const bool more = __syncthreads_or(changed) != 0;
        if (!more) break;
        if (passes >= cap) { unfinished = true; break; }
    }

    // Face predictions: would the neighbour's cell across face s improve by >= eps?  Its
    // neighbour further out is not staged and counts as +INF.
    if (tid < 6) {
        s_u32[tid] = 0u;
        s_f32[tid] = __builtin_inff();
    }
    __syncthreads();
    for (int j = tid; j < 6 * B * B; j += kT3Threads) {
        const int s = j / (B * B), f = j % (B * B);
        const int u = f % B, v = f / B;
        int dx, dy, dz;
        side_step3(s, dx, dy, dz);
        const int ntx = tx + dx, nty = ty + dy, ntz = tz + dz;
        if (ntx < 0 || nty < 0 || ntz < 0 || ntx >= (int)in.tiles_x || nty >= (int)in.tiles_y ||
            ntz >= (int)in.tiles_z)
            continue;
        // halo cell (bx, by, bz) in box coordinates
        int bx, by, bz;
        if (s == 0 || s == 2)      { bx = s == 0 ? 0 : B + 1; by = u + 1; bz = v + 1; }
        else if (s == 1 || s == 3) { by = s == 1 ? 0 : B + 1; bx = u + 1; bz = v + 1; }
        else                       { bz = s == 4 ? 0 : B + 1; bx = u + 1; by = v + 1; }
        const int gx = ox + bx - 1, gy = oy + by - 1, gz = oz + bz - 1;
        if (gx >= W || gy >= H || gz >= D) continue;
        const int c = (bz * PB + by) * PB + bx;
        float ax, ay, az;
        if (s == 0 || s == 2) ax = t_s[c + (s == 0 ? 1 : -1)];
        else ax = fminf(t_s[c - 1], t_s[c + 1]);
        if (s == 1 || s == 3) ay = t_s[c + (s == 1 ? PB : -PB)];
        else ay = fminf(t_s[c - PB], t_s[c + PB]);
        if (s == 4 || s == 5) az = t_s[c + (s == 4 ? PB2 : -PB2)];
        else az = fminf(t_s[c - PB2], t_s[c + PB2]);
        const float t0 = t_s[c];
        const float tn = godunov3(ax, ay, az, c_s[c]);
        if (tn < t0 && t0 - tn >= eps) {
            atomicOr(&s_u32[s], 1u);
            atomicMin(reinterpret_cast<uint32_t*>(s_f32) + s, __float_as_uint(tn));
        }
    }
    __syncthreads();
    if (tid < 6 && s_u32[tid]) {
        int dx, dy, dz;
        side_step3(tid, dx, dy, dz);
        const uint32_t nb = in.tile_base + (uint32_t)(tx + dx) +
                            in.tiles_x * ((uint32_t)(ty + dy) + in.tiles_y * (uint32_t)(tz + dz));
        activate(a, nb, wave, nxt, p, kMaskOpenAll, s_f32[tid]);
    }

    // Commit improved cells (one writer per tile per wave), tile minimum.
    float tmin = __builtin_inff();
    uint32_t improved = 0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int i = tid + kT3Threads * k;
        const int x = i % B, y = (i / B) % B, z = i / (B * B);
        const float tn = t_s[q[k]];
        if (ox + x < W && oy + y < H && oz + z < D) {
            tmin = fminf(tmin, tn);
            if (tn < told[k]) {
                const size_t idx = ((size_t)(oz + z) * (size_t)H + (size_t)(oy + y)) * (size_t)W + (size_t)(ox + x);
                gmem(in.time)[idx] = tn;
                ++improved;
            }
        }
    }
    tmin = wave_min(tmin);
    improved = wave_sum(improved);
    const int w = tid >> 6, L = tid & 63;
    if (L == 0) {
        s_f32[8 + w] = tmin;
        s_u32[8 + w] = improved;
    }
    __syncthreads();
    if (tid == 0) {
        const float tm = fminf(fminf(s_f32[8], s_f32[9]), fminf(s_f32[10], s_f32[11]));
        a.tmin[g] = tm;
        stat_add(&a.tstat[g], 1u);
        stat_add(&a.tstat[(size_t)a.tiles + g], passes);
        stat_add(&a.tstat[2 * (size_t)a.tiles + g], s_u32[8] + s_u32[9] + s_u32[10] + s_u32[11]);
        if (unfinished) activate(a, g, wave, nxt, p, kMaskOpenAll, tm);
    }
}

template <int KIND>
__global__ __launch_bounds__(kT3Threads) __attribute__((amdgpu_waves_per_eu(8))) void k_solve3d(SolveArgs a) {
    __shared__ float t_s[kT3Box * kT3Box * kT3Box];
    __shared__ float c_s[kT3Box * kT3Box * kT3Box];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[16];
    __shared__ uint32_t s_prefix[kShards + 1];
    const int tid = threadIdx.x;
    const GhcmInst in0 = single_inst(a);
    const GhcmCtrl* ctrl = a.ctrl;
    publish_progress(a, ctrl);
    const uint32_t done = ctrl->done;
    const uint32_t nj = ctrl->n_jobs;
    const uint32_t wave = ctrl->wave;
    const uint32_t cur = ctrl->cur;
    const uint32_t from_list = ctrl->jobs_from_list;
    const uint32_t ncarry = ctrl->n_carry;
    // Every thread reads a valid word (no branch around the load).
    const uint32_t pv = ctrl->prefix[min(tid, (int)kShards)];
    // Work-groups past the jobs still take their share of the handed-off carry.
    // A compiler barrier: every load above is issued before the done test (without it the
    // short-circuit test lets the compiler sink the later loads below it, one memory round
    // trip each).
    __asm__ volatile("" ::: "memory");
    if (done || (blockIdx.x >= nj && ncarry == 0u)) return;
    const uint32_t p = wave & 1u;
    if (tid <= (int)kShards) s_prefix[tid] = pv;
    __syncthreads();
    for (uint32_t j = blockIdx.x; j < nj; j += gridDim.x) {
        const uint32_t g = job_tile(a, s_prefix, cur, from_list, j);
        if (tid == 0) claim_tile(a, g, p);
        solve3d_tile<KIND>(a, g, wave, cur ^ 1u, p, t_s, c_s, s_u32, s_f32, in0);
        __syncthreads();
    }
    if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
}

// ---------------------------------------------------------------------------------------
// Three-dimensional sweep tile solver: eight wave64s per work-group relax one 64 x 8 x 8 tile
// with the eight Gauss-Seidel orders of 3-D fast sweeping (Zhao 2005), one octant per wave,
// concurrently on the same LDS box.  Lane L owns the row (y, z) = (L mod 8, L / 8) and walks
// along x (towards +x or -x with the octant); the octant's rank of the row, r = y' + z' (y'
// and z' counted in the octant's directions, 0..14), skews the walk: at step s the lane is at
// x-position (s - r) mod 64.  The rows upwind in y and z have rank r - 1, so they relaxed the
// same x one step earlier: their results arrive through a DPP row shift (y, the neighbouring
// lane) and ds_bpermute (z, eight lanes away), the upwind x value is the lane's own previous
// result, and each is merged with an LDS read of that cell (improvements of the other seven
// sweeps).  64 + 14 steps give every row one complete ordered walk, so a round of the eight
// sweeps is one full fast-sweeping iteration over the tile in 78 steps; before and after its
// walk a lane relaxes cells of the neighbouring walks instead of idling (every value involved
// is an upper bound).  Rounds repeat until no cell improves by >= eps or the round cap.
// ---------------------------------------------------------------------------------------
constexpr int kS3X = 64, kS3Y = 8, kS3Z = 8;       // tile extents
constexpr int kS3Threads = 512;                    // eight wave64s, one octant each
constexpr int kS3PX = kS3X + 2;                    // box row: x = -1 .. 64
// Box plane stride in words (>= 10 rows of 66).  A step's 64 lanes address the cells
// (z, y, x) = (z, y, c -+ y -+ z); with rows of 66 words and planes of 695 the 64 addresses
// fall in 64 different banks for half of the octants and at most two per bank for the others
// (no stride from 660 to 800 does better for all eight).
constexpr int kS3Plane = 695;
constexpr int kS3Box = (kS3Z + 2) * kS3Plane;      // words of the time box; the cost box follows
constexpr int kS3MaxRank = (kS3Y - 1) + (kS3Z - 1);
constexpr uint32_t kS3HardRounds = 64;
static_assert(kS3Y == 8 && kS3Z == 8, "lane L owns row (L mod 8, L / 8)");

__device__ __forceinline__ int s3_box(int bx, int by, int bz) { return bz * kS3Plane + by * kS3PX + bx; }

// One round of octant (DX, SY, SZ) over the LDS box (time box at shb, cost box kS3Box words
// after it), run by one wave64.
template <int DX, int SY, int SZ>
__device__ __forceinline__ void sweep3_round(char* const shb, const int L) {
    constexpr int PX = kS3PX, PL = kS3Plane;
    // Word offsets from the corner (x-1, y-1, z-1) of the cell's 3 x 3 x 3 neighbourhood: every
    // read and the merge are immediate offsets from one per-step address.
    constexpr int kCell = PL + PX + 1;
    constexpr int kUpX = kCell - DX;
    constexpr int kUpY = kCell - SY * PX;
    constexpr int kUpZ = kCell - SZ * PL;
    constexpr int kCost = kS3Box + kCell;
    // Rows whose upwind z neighbour is in the halo: z = 0 (7) for SZ > 0 (< 0).
    constexpr uint64_t kZHalo = SZ > 0 ? 0x00000000000000FFull : 0xFF00000000000000ull;
    // DPP row shift by one lane towards higher lanes (SY > 0: lane L reads lane L-1) or lower.
    // The rows whose upwind y neighbour is in the halo (y = 0 for SY > 0, 7 for SY < 0) either
    // have no source lane in their row of 16 or take the row y = 7 (0) of the neighbouring z:
    // those rows (kYLast) hand on +INF instead of their result, so every halo row's y value is
    // its LDS read alone.
    constexpr int kDppY = SY > 0 ? 0x111 : 0x101;
    constexpr uint64_t kYLast = SY > 0 ? 0x8080808080808080ull : 0x0101010101010101ull;
    const int y = L & 7, z = L >> 3;
    const int rank = (SY > 0 ? y : kS3Y - 1 - y) + (SZ > 0 ? z : kS3Z - 1 - z);
    const int zsrc = 4 * ((L - 8 * SZ) & 63);           // ds_bpermute byte address of the z-upwind lane
    const int x0 = DX > 0 ? 0 : kS3X - 1;
    const int rs = 4 * (z * PL + y * PX + x0);           // the walk's first cell (corner-relative)
    // The first load steps to x-position (0 - rank) mod 64, or to rs where that is 0.
    int bb = rs + 4 * DX * (kS3X - 1 - rank);
    // 78 steps complete every walk; the loop runs 80 (four per iteration, never unrolled further).
    constexpr int S = kS3X + kS3MaxRank;
    constexpr int SE = (S + 3) & ~3;
    struct Nb {
        float ux, uy, uz, cc;
        int b;
        bool first;
    };
    // Single-word loads at immediate offsets from the step's address (relaxed atomic loads are
    // not paired into ds_read2, which would need a second address register per step).
    auto at = [&](int b, int word) -> float {
        return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(shb + b + 4 * word),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    };
    // Step s: lanes with rank == s mod 64 are at x-position 0 (their walk restarts at rs).  The
    // test for the next step is made one load ahead, so the select does not wait on the
    // compare's lane mask.
    bool first_next = rank == 0;
    auto load = [&](Nb& o, int st) {
        o.first = first_next;
        bb = o.first ? rs : bb + 4 * DX;
        o.b = bb;
        o.ux = at(bb, kUpX);
        o.uy = at(bb, kUpY);
        o.uz = at(bb, kUpZ);
        o.cc = at(bb, kCost);
        first_next = rank == ((st + 1) & 63);
    };
    auto umin = [](float u, float v) {
        return __uint_as_float(min(__float_as_uint(u), __float_as_uint(v)));
    };
    const float inf = __builtin_inff();
    // The previous step's result of this lane (upwind x), of the y-upwind lane (a DPP row shift
    // folded into the minimum: its `old` is the minimum's identity and its bound control is
    // off, so a lane without a source lane keeps its LDS read) and of the z-upwind lane
    // (ds_bpermute, issued as soon as the result exists, ahead of the next step's LDS reads).
    float prev = inf, ysend = inf, zv = inf;
    auto step = [&](const Nb& n) {
        const float ax = umin(n.ux, n.first ? inf : prev);
        const float ay = __uint_as_float(
            min(__float_as_uint(n.uy),
                (uint32_t)__builtin_amdgcn_update_dpp(-1, __float_as_int(ysend), kDppY, 0xF, 0xF, false)));
        const float az = umin(n.uz, __builtin_amdgcn_inverse_ballot_w64(kZHalo) ? inf : zv);
        const float t = godunov3(ax, ay, az, n.cc);
        __hip_atomic_fetch_min(reinterpret_cast<uint32_t*>(shb + n.b + 4 * kCell), __float_as_uint(t),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        prev = t;
        zv = __int_as_float(__builtin_amdgcn_ds_bpermute(zsrc, __float_as_int(t)));
        ysend = __builtin_amdgcn_inverse_ballot_w64(kYLast) ? inf : t;
    };
    // Loads of step s+1 are issued before step s is computed (two register sets): their cells
    // are upper bounds whatever the interleaving, and the values step s produces reach step
    // s+1 through registers.
    Nb A, B;
    load(A, 0);
#pragma clang loop unroll_count(2)
    for (int st = 0; st < SE; st += 2) {
        load(B, st + 1);
        __builtin_amdgcn_sched_barrier(0);
        step(A);
        __builtin_amdgcn_sched_barrier(0);
        load(A, st + 2);
        __builtin_amdgcn_sched_barrier(0);
        step(B);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Stage the (kS3X + 2) x 10 x 10 box of 3-D sweep tile g at (ox, oy, oz): the time box, the
// cost box at the tile and its six faces (edges and corners are never read).  Rows of the box
// along x are read as aligned runs of 64 cells (the tile's x range), one run per wave-uniform
// row job.  The x = -1 and x = 64 words of the tile's 64 rows are the neighbours' edge columns
// (a.edge, a.cedge: 64 contiguous floats per side, row r = (r mod 8, r / 8)); those of the 36
// halo rows (time only) are read one per thread.  Out-of-grid cells read +INF.  Every global
// load is issued before the first LDS store.
template <int KIND>
__device__ __forceinline__ void stage_sweep3(const SolveArgs& a, const GhcmInst& in, uint32_t g, int tx,
                                             int ox, int oy, int oz, int tid, float* t_s, float* c_s) {
    constexpr int kRowJobs = 64 + 64 + 36 + 32;        // time, cost interior; time, cost halo rows
    constexpr int NR = (kRowJobs + 7) / 8;
    constexpr int kHaloWords = 4 * 64 + 2 * 36;        // edge columns (time, cost); halo rows' words
    // The wave index stays a vector value here: as a scalar, the 25 row jobs' addresses
    // exceed the scalar registers and spill.
    const int w = tid >> 6, L = tid & 63;
    const int W = (int)in.w, H = (int)in.h, D = (int)in.d;
    const char* const tb = reinterpret_cast<const char*>(in.time);
    const char* const fb = static_cast<const char*>(in.field);
    constexpr int fbytes = field_bytes<KIND>();
    // Row job q -> field and box row (by, bz); false past the last job.  Jobs 0-63: time
    // rows of the tile; 64-127: its cost rows; 128-163: time halo rows (the y-halo rows of the
    // eight tile planes, then both z-halo planes whole); 164-195: cost halo rows (the same
    // face rows without the box edges).
    auto row_job = [](int q, bool& cost, int& by, int& bz) -> bool {
        if (q < 128) {
            cost = q >= 64;
            const int r = q & 63;
            by = (r & 7) + 1;
            bz = (r >> 3) + 1;
            return true;
        }
        q -= 128;
        if (q < 36) {
            cost = false;
            if (q < 16) { by = (q & 1) ? kS3Y + 1 : 0; bz = (q >> 1) + 1; }
            else        { q -= 16; bz = q >= 10 ? kS3Z + 1 : 0; by = q % 10; }
            return true;
        }
        q -= 36;
        if (q < 32) {
            cost = true;
            if (q < 16) { by = (q & 1) ? kS3Y + 1 : 0; bz = (q >> 1) + 1; }
            else        { q -= 16; bz = q >= 8 ? kS3Z + 1 : 0; by = (q & 7) + 1; }
            return true;
        }
        return false;
    };
    float v[NR];
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        bool cost = false;
        int by = 0, bz = 0;
        v[m] = __builtin_inff();
        if (row_job(w + 8 * m, cost, by, bz)) {
            const int gx = ox + L, gy = oy + by - 1, gz = oz + bz - 1;
            if (gx < W && gy >= 0 && gz >= 0 && gy < H && gz < D) {
                const size_t idx = ((size_t)gz * (size_t)H + (size_t)gy) * (size_t)W + (size_t)gx;
                v[m] = cost ? load_cost_at<KIND>(fb + idx * fbytes)
                            : *gmem(reinterpret_cast<const float*>(tb + idx * 4));
            }
        }
    }
    float hv = __builtin_inff();
    int hq = -1;
    bool hcost = false;
    if (tid < 4 * 64) {
        // Edge columns: tid 0-127 time, 128-255 cost; the left halo column is the left
        // neighbour's last column, the right one the right neighbour's first column.
        const int r = tid & 63, right = (tid >> 6) & 1;
        hcost = tid >= 128;
        const int y = r & 7, z = r >> 3;
        hq = s3_box(right ? kS3X + 1 : 0, y + 1, z + 1);
        const bool has = right ? ox + kS3X < W : tx > 0;
        if (has && oy + y < H && oz + z < D) {
            const size_t o = (right ? (size_t)g + 1 : (size_t)g - 1) * 2 * 64 + (right ? 0 : 64) + (size_t)r;
            hv = *gmem(hcost ? a.cedge + o : a.edge + o);
        }
    } else if (tid < kHaloWords) {
        // The x = -1 and x = 64 time words of the 36 halo rows (the y-halo rows of the tile
        // planes, then both z-halo planes whole).
        int q = (tid - 4 * 64) >> 1, by, bz;
        if (q < 16) { by = (q & 1) ? kS3Y + 1 : 0; bz = (q >> 1) + 1; }
        else        { q -= 16; bz = q >= 10 ? kS3Z + 1 : 0; by = q % 10; }
        const int bx = (tid & 1) ? kS3X + 1 : 0;
        hq = s3_box(bx, by, bz);
        const int gx = ox + bx - 1, gy = oy + by - 1, gz = oz + bz - 1;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < W && gy < H && gz < D) {
            const size_t idx = ((size_t)gz * (size_t)H + (size_t)gy) * (size_t)W + (size_t)gx;
            hv = *gmem(reinterpret_cast<const float*>(tb + idx * 4));
        }
    }
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        bool cost = false;
        int by = 0, bz = 0;
        if (row_job(w + 8 * m, cost, by, bz)) (cost ? c_s : t_s)[s3_box(L + 1, by, bz)] = v[m];
    }
    if (hq >= 0) (hcost ? c_s : t_s)[hq] = hv;
}

// One visit of 3-D sweep tile g: stage, rounds of the eight octant sweeps, commit, face
// predictions for the six neighbours, activations.  Thread (wave w, lane L) owns the cells
// (x = L, y = w, z = 0..7) for the convergence check and the commit.
template <int KIND>
__device__ void sweep3_tile(const SolveArgs& a, uint32_t g, uint32_t wave, uint32_t nxt, uint32_t p,
                            float* sh, uint32_t* s_u32, float* s_f32, float* s_fk, GhcmInst in_single) {
    constexpr int PX = kS3PX, PL = kS3Plane;
    float* const t_s = sh;
    float* const c_s = sh + kS3Box;
    const int tid = threadIdx.x;
    // The wave index, scalar: the octant, the face blocks and the owned row are wave-uniform.
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), L = tid & 63;
    GhcmInst in = in_single;
    if (a.n_inst != 1) in = uniform_inst(a.inst[a.tinst[g]]);
    const uint32_t lt = g - in.tile_base;
    const int tx = (int)(lt % in.tiles_x);
    const int ty = (int)((lt / in.tiles_x) % in.tiles_y);
    const int tz = (int)(lt / (in.tiles_x * in.tiles_y));
    const int W = (int)in.w, H = (int)in.h, D = (int)in.d;
    const int ox = tx * kS3X, oy = ty * kS3Y, oz = tz * kS3Z;
    const int sx = min(kS3X, W - ox), sy = min(kS3Y, H - oy), sz = min(kS3Z, D - oz);
    // Face keys of this visit (+INF: the neighbour across the face is not activated); read
    // by the same threads at the end of the previous visit, written by atomics after the
    // staging barrier.
    if (tid < 6) s_fk[tid] = __builtin_inff();
    stage_sweep3<KIND>(a, in, g, tx, ox, oy, oz, tid, t_s, c_s);
    __syncthreads();
    float told[kS3Z], snap[kS3Z];
#pragma unroll
    for (int k = 0; k < kS3Z; ++k) {
        told[k] = t_s[s3_box(L + 1, w + 1, k + 1)];
        snap[k] = told[k];
    }
    const float eps = a.eps;
    const uint32_t cap = (a.iter_cap == 0u || a.iter_cap > kS3HardRounds) ? kS3HardRounds : a.iter_cap;
    char* const shb = reinterpret_cast<char*>(sh);
    uint32_t rounds = 0;
    bool unfinished = false;
    for (;;) {
        // Wave w runs octant w: bit 0 -x, bit 1 -y, bit 2 -z (the branch is wave-uniform).
        switch (w) {
            case 0: sweep3_round<1, 1, 1>(shb, L); break;
            case 1: sweep3_round<-1, 1, 1>(shb, L); break;
            case 2: sweep3_round<1, -1, 1>(shb, L); break;
            case 3: sweep3_round<-1, -1, 1>(shb, L); break;
            case 4: sweep3_round<1, 1, -1>(shb, L); break;
            case 5: sweep3_round<-1, 1, -1>(shb, L); break;
            case 6: sweep3_round<1, -1, -1>(shb, L); break;
            default: sweep3_round<-1, -1, -1>(shb, L); break;
        }
        ++rounds;
        __syncthreads();
        // With one round per visit its change test is made in the commit pass below, which
        // reads the same cells (the round began at the staged times, `told`).
        if (cap == 1u) break;
        bool chg = false;
#pragma unroll
        for (int k = 0; k < kS3Z; ++k) {
            const float now = t_s[s3_box(L + 1, w + 1, k + 1)];
            const float d = snap[k] - now;          // NaN (INF - INF) counts as no change
            if (d > 0.0f && d >= eps) chg = true;
            snap[k] = now;
        }
        const bool any = __ballot(chg) != 0ull;
        if (L == 0) s_u32[w] = any ? 1u : 0u;
        __syncthreads();
        // No third barrier: a wave writes s_u32[w] again only after the next round's first
        // barrier, which every wave reaches after reading it here.
        uint32_t more = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) more |= s_u32[i];
        if (!more) break;
        if (rounds >= cap) { unfinished = true; break; }
    }
    // Commit improved cells (one writer per tile per wave): wave w writes row y = w of every
    // z-plane, 64 contiguous cells each.  With one round per visit, whether it changed the
    // tile (then the tile re-queues itself) is tested here against `told` (cells outside the
    // grid are +INF and never change); each wave's answer goes to s_u32[w], which that path
    // leaves unused, and is read after the barrier before the activations.
    const bool one_round = cap == 1u;
    float tmin = __builtin_inff();
    uint32_t improved = 0;
    bool last_chg = false;
#pragma unroll
    for (int k = 0; k < kS3Z; ++k) {
        if (L < sx && w < sy && k < sz) {
            const float tn = t_s[s3_box(L + 1, w + 1, k + 1)];
            tmin = fminf(tmin, tn);
            if (tn < told[k]) {
                const size_t idx = ((size_t)(oz + k) * (size_t)H + (size_t)(oy + w)) * (size_t)W + (size_t)(ox + L);
                gmem(in.time)[idx] = tn;
                ++improved;
                if (told[k] - tn >= eps) last_chg = true;
            }
        }
    }
    if (one_round) {
        // Only on this path: after a round check other waves may still be reading s_u32[0..7].
        const bool wave_chg = __ballot(last_chg) != 0ull;
        if (L == 0) s_u32[w] = wave_chg ? 1u : 0u;
    }
    // Edge columns from the final box: wave 0 the first column, wave 1 the last one when the
    // tile is full width (only then has it a right neighbour that reads it); lane L is row
    // (L mod 8, L / 8), and rows past the grid are never read and not written.
    if (w < 2 && (w == 0 || sx == kS3X)) {
        const int y = L & 7, z = L >> 3;
        if (y < sy && z < sz)
            gmem(a.edge)[(size_t)g * 2 * 64 + (size_t)w * 64 + L] = t_s[s3_box(w == 0 ? 1 : kS3X, y + 1, z + 1)];
    }
    // Face predictions: would the neighbour's cell across face s improve by >= eps from this
    // tile's values (data/eikonal_ghcm_systolic.cl:8-41 in three dimensions)?  The 2176 face
    // cells form 34 blocks of 64 (x-, x+, 8 of y-, 8 of y+, 8 of z-, 8 of z+); wave w takes
    // blocks w, w+8, ..., so a block's face is wave-uniform.  The halo cell's neighbour further
    // out is not staged and counts as +INF.
    // One code path for every face: the block's face gives (wave-uniform) the halo cell's box
    // word as base + u * su + v * sv for in-face coordinates (u, v) = (f & mu, f >> shv), the
    // word offset of its one staged neighbour inside the tile, and the offsets of its two
    // lateral pairs; godunov3 sorts its three axis values, so their order is free.
    const int ntx = (int)in.tiles_x, nty = (int)in.tiles_y, ntz = (int)in.tiles_z;
    auto umin = [](float u, float v) { return __uint_as_float(min(__float_as_uint(u), __float_as_uint(v))); };
    for (int bi = w; bi < 34; bi += 8) {
        int s, blk;
        if (bi < 2)       { s = bi == 0 ? 0 : 2; blk = 0; }
        else if (bi < 18) { s = bi < 10 ? 1 : 3; blk = (bi - 2) & 7; }
        else              { s = bi < 26 ? 4 : 5; blk = (bi - 18) & 7; }
        const bool has_nb = s == 0 ? tx > 0 : s == 2 ? tx + 1 < ntx : s == 1 ? ty > 0
                          : s == 3 ? ty + 1 < nty : s == 4 ? tz > 0 : tz + 1 < ntz;
        if (!has_nb) continue;
        const int f = blk * 64 + L;
        const bool xf = s == 0 || s == 2, yf = s == 1 || s == 3;
        const int mu = xf ? 7 : 63, shv = xf ? 3 : 6;
        const int su = xf ? PX : 1, sv = xf ? PL : yf ? PL : PX;
        const int eu = xf ? sy : sx, ev = xf ? sz : yf ? sz : sy;
        // the face plane's box coordinate along its normal, and the unit step inwards
        const int nrm = s == 0 ? 0 : s == 2 ? kS3X + 1 : s == 1 ? 0 : s == 3 ? kS3Y + 1 : s == 4 ? 0 : kS3Z + 1;
        const int nst = xf ? 1 : yf ? PX : PL;
        const int base = xf ? nrm + PX + PL : yf ? 1 + nrm * PX + PL : 1 + PX + nrm * PL;
        const int in_off = (s == 0 || s == 1 || s == 4) ? nst : -nst;
        const int l1 = xf ? PX : 1, l2 = xf ? PL : yf ? PL : PX;
        const int u = f & mu, v = f >> shv;
        float key = __builtin_inff();
        if (u < eu && v < ev) {
            const int c = base + u * su + v * sv;
            const float a_in = t_s[c + in_off];
            const float a_1 = umin(t_s[c - l1], t_s[c + l1]);
            const float a_2 = umin(t_s[c - l2], t_s[c + l2]);
            const float t0 = t_s[c];
            const float tn = godunov3(a_in, a_1, a_2, c_s[c]);
            if (tn < t0 && t0 - tn >= eps) key = tn;
        }
        // Most blocks predict no improvement: only the others reduce their key and merge it.
        if (__ballot(key < __builtin_inff()) == 0ull) continue;
        key = wave_min(key);
        if (L == 0)
            __hip_atomic_fetch_min(reinterpret_cast<uint32_t*>(s_fk) + s, __float_as_uint(key),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    tmin = wave_min(tmin);
    improved = wave_sum(improved);
    if (L == 0) {
        s_f32[w] = tmin;
        s_u32[8 + w] = improved;
    }
    __syncthreads();
    float tm = s_f32[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) tm = fminf(tm, s_f32[i]);
    if (one_round) {
        uint32_t any = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) any |= s_u32[i];
        unfinished = any != 0u;
    }
    // Every activation after one barrier: the faces' neighbours (lanes 0-5 of wave 0) and,
    // when the round cap stopped the tile, the tile itself (wave 1).
    if (tid < 6) {
        const float k = s_fk[tid];
        if (k < __builtin_inff()) {
            int dx, dy, dz;
            side_step3(tid, dx, dy, dz);
            const uint32_t nb = in.tile_base + (uint32_t)(tx + dx) +
                                in.tiles_x * ((uint32_t)(ty + dy) + in.tiles_y * (uint32_t)(tz + dz));
            activate(a, nb, wave, nxt, p, kMaskOpenAll, k);
        }
    }
    if (tid == 64 && unfinished) activate(a, g, wave, nxt, p, kMaskOpenAll, tm);
    if (tid == 128) {
        uint32_t imp = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) imp += s_u32[8 + i];
        a.tmin[g] = tm;
        stat_add(&a.tstat[g], 1u);
        stat_add(&a.tstat[(size_t)a.tiles + g], rounds);
        stat_add(&a.tstat[2 * (size_t)a.tiles + g], imp);
    }
}

// 2 work-groups per CU (the LDS boxes take 53.8 KB each): at most 128 VGPRs.
template <int KIND>
__global__ __launch_bounds__(kS3Threads) __attribute__((amdgpu_waves_per_eu(4))) void k_solve3d_sweep(SolveArgs a) {
    __shared__ float sh[2 * kS3Box];
    __shared__ uint32_t s_u32[16];
    __shared__ float s_f32[8];
    __shared__ float s_fk[8];
    __shared__ uint32_t s_prefix[kShards + 1];
    const int tid = threadIdx.x;
    const GhcmInst in0 = single_inst(a);
    const GhcmCtrl* ctrl = a.ctrl;
    publish_progress(a, ctrl);
    const uint32_t done = ctrl->done;
    const uint32_t nj = ctrl->n_jobs;
    const uint32_t wave = ctrl->wave;
    const uint32_t cur = ctrl->cur;
    const uint32_t from_list = ctrl->jobs_from_list;
    const uint32_t ncarry = ctrl->n_carry;
    // Every thread reads a valid word (no branch around the load).
    const uint32_t pv = ctrl->prefix[min(tid, (int)kShards)];
    // Work-groups past the jobs still take their share of the handed-off carry.
    // A compiler barrier: every load above is issued before the done test (without it the
    // short-circuit test lets the compiler sink the later loads below it, one memory round
    // trip each).
    __asm__ volatile("" ::: "memory");
    if (done || (blockIdx.x >= nj && ncarry == 0u)) return;
    const uint32_t p = wave & 1u;
    if (tid <= (int)kShards) s_prefix[tid] = pv;
    __syncthreads();
    for (uint32_t j = blockIdx.x; j < nj; j += gridDim.x) {
        const uint32_t g = job_tile(a, s_prefix, cur, from_list, j);
        if (tid == 0) claim_tile(a, g, p);
        sweep3_tile<KIND>(a, g, wave, cur ^ 1u, p, sh, s_u32, s_f32, s_fk, in0);
        __syncthreads();
    }
    if (ncarry) carry_handed_off(a, nj, ncarry, wave, cur ^ 1u, p);
}

// Seeds of 3-D solves: (x, y, z, instance) quadruples; T = 0 at the seed, and the tiles of
// the seed and of its six neighbours are queued for the first wave with every cell open.
__global__ void k_seed3d(SolveArgs a, const uint32_t* seeds, uint32_t n_seeds, uint32_t tw, uint32_t th,
                         uint32_t td) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seeds) return;
    const uint32_t x = seeds[4 * s], y = seeds[4 * s + 1], z = seeds[4 * s + 2], ii = seeds[4 * s + 3];
    const GhcmInst in = a.inst[ii];
    gmem(in.time)[((size_t)z * in.h + y) * in.w + x] = 0.0f;
    if (a.edge) {
        // A seed on its tile's first or last column also sets that tile's edge column.
        const uint32_t t = in.tile_base + x / tw + in.tiles_x * (y / th + in.tiles_y * (z / td));
        const uint32_t eh = th * td, r = y % th + th * (z % td), cx = x % tw;
        if (cx == 0) gmem(a.edge)[(size_t)t * 2 * eh + r] = 0.0f;
        if (cx == tw - 1) gmem(a.edge)[(size_t)t * 2 * eh + eh + r] = 0.0f;
    }
    for (int d = 0; d < 7; ++d) {
        int dx = 0, dy = 0, dz = 0;
        if (d > 0) side_step3(d - 1, dx, dy, dz);
        const int px = (int)x + dx, py = (int)y + dy, pz = (int)z + dz;
        if (px < 0 || py < 0 || pz < 0 || px >= (int)in.w || py >= (int)in.h || pz >= (int)in.d) continue;
        const uint32_t t = in.tile_base + (uint32_t)px / tw +
                           in.tiles_x * ((uint32_t)py / th + in.tiles_y * ((uint32_t)pz / td));
        atomicMin(&a.key[t], 0u);
        atomicOr(&a.mask[t], kMaskOpenAll);
        const uint32_t old = atomicMax(&a.qwave[t], 1u);
        if (old < 1u) list_append(a, 0u, t);
    }
}

// ---- launcher table ----
// ev0/ev1 (optional) are stamped by the dispatch itself (hipExtLaunchKernelGGL), so
// timing a launch adds no marker packets to the stream.
template <int TW, int TH, int KIND, int MODE, bool HOST>
static hipError_t launch_solve_t(const SolveArgs& a, uint32_t grid, hipStream_t s,
                                 hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve<TW, TH, KIND, MODE, HOST>), dim3(grid), dim3(64), 0, s,
                              ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL((k_solve<TW, TH, KIND, MODE, HOST>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <int TW, int TH, int KIND, int MODE, bool HOST>
static int occ_solve_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve<TW, TH, KIND, MODE, HOST>, 64, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int TW, int TH, int KIND, bool HOST, int WPE = 1>
static int occ_sweep_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve_sweep<TW, TH, KIND, HOST, WPE>,
                                                     kSweepThreads, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int TW, int TH, int KIND, bool HOST, int WPE = 1>
static hipError_t launch_sweep_t(const SolveArgs& a, uint32_t grid, hipStream_t s,
                                 hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve_sweep<TW, TH, KIND, HOST, WPE>), dim3(grid),
                              dim3(kSweepThreads), 0, s, ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL((k_solve_sweep<TW, TH, KIND, HOST, WPE>), dim3(grid),
                           dim3(kSweepThreads), 0, s, a);
    return hipGetLastError();
}

template <int TW, int TH, int KIND>
static hipError_t launch_fused_t(const SolveArgs& a, uint32_t grid, hipStream_t s, hipEvent_t ev0,
                                 hipEvent_t ev1) {
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve_sweep_fused<TW, TH, KIND>), dim3(grid), dim3(kSweepThreads),
                              0, s, ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL((k_solve_sweep_fused<TW, TH, KIND>), dim3(grid), dim3(kSweepThreads), 0,
                           s, a);
    return hipGetLastError();
}

template <int TW, int TH, int KIND>
static int occ_fused_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve_sweep_fused<TW, TH, KIND>,
                                                     kSweepThreads, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int TW, int TH>
static SolveKernel pick_fused(int kind) {
    switch (kind) {
        case 0: return {&launch_fused_t<TW, TH, 0>, &occ_fused_t<TW, TH, 0>};
        case 1: return {&launch_fused_t<TW, TH, 1>, &occ_fused_t<TW, TH, 1>};
        case 2: return {&launch_fused_t<TW, TH, 2>, &occ_fused_t<TW, TH, 2>};
        case 3: return {&launch_fused_t<TW, TH, 3>, &occ_fused_t<TW, TH, 3>};
    }
    return {nullptr, nullptr};
}

SolveKernel get_fused_kernel(int tw, int th, int kind, int mode) {
    if (mode != 2) return {nullptr, nullptr};
    if (tw == 64 && th == 64) return pick_fused<64, 64>(kind);
    if (tw == 64 && th == 32) return pick_fused<64, 32>(kind);
    if (tw == 32 && th == 32) return pick_fused<32, 32>(kind);
    return {nullptr, nullptr};
}

template <int TW, int TH, int KIND>
static hipError_t launch_persistent_t(const SolveArgs& a, uint32_t grid, hipStream_t s) {
    void* args[] = {const_cast<SolveArgs*>(&a)};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_persistent_sweep<TW, TH, KIND>),
                                      dim3(grid), dim3(kSweepThreads), args, 0, s);
}

template <int TW, int TH, int KIND>
static int occ_persistent_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_persistent_sweep<TW, TH, KIND>,
                                                     kSweepThreads, 0) != hipSuccess)
        return 0;
    return nb;
}

template <int TW, int TH>
static PersistentKernel pick_persistent(int kind) {
    switch (kind) {
        case 0: return {&launch_persistent_t<TW, TH, 0>, &occ_persistent_t<TW, TH, 0>};
        case 1: return {&launch_persistent_t<TW, TH, 1>, &occ_persistent_t<TW, TH, 1>};
        case 2: return {&launch_persistent_t<TW, TH, 2>, &occ_persistent_t<TW, TH, 2>};
        case 3: return {&launch_persistent_t<TW, TH, 3>, &occ_persistent_t<TW, TH, 3>};
    }
    return {nullptr, nullptr};
}

PersistentKernel get_persistent_kernel(int tw, int th, int kind, int mode) {
    if (mode != 2) return {nullptr, nullptr};
    if (tw == 64 && th == 64) return pick_persistent<64, 64>(kind);
    if (tw == 64 && th == 32) return pick_persistent<64, 32>(kind);
    if (tw == 64 && th == 16) return pick_persistent<64, 16>(kind);
    if (tw == 32 && th == 32) return pick_persistent<32, 32>(kind);
    return {nullptr, nullptr};
}

template <int TW, int TH, bool HOST>
static SolveKernel pick_sweep_kind(int kind) {
    switch (kind) {
        case 0: return {&launch_sweep_t<TW, TH, 0, HOST>, &occ_sweep_t<TW, TH, 0, HOST>};
        case 1: return {&launch_sweep_t<TW, TH, 1, HOST>, &occ_sweep_t<TW, TH, 1, HOST>};
        case 2: return {&launch_sweep_t<TW, TH, 2, HOST>, &occ_sweep_t<TW, TH, 2, HOST>};
        case 3: return {&launch_sweep_t<TW, TH, 3, HOST>, &occ_sweep_t<TW, TH, 3, HOST>};
    }
    return {nullptr, nullptr};
}

template <int TW, int TH>
static SolveKernel pick_sweep(int kind, bool host) {
    return host ? pick_sweep_kind<TW, TH, true>(kind) : pick_sweep_kind<TW, TH, false>(kind);
}

template <int TW, int TH, int MODE, bool HOST>
static SolveKernel pick_kind(int kind) {
    switch (kind) {
        case 0: return {&launch_solve_t<TW, TH, 0, MODE, HOST>, &occ_solve_t<TW, TH, 0, MODE, HOST>};
        case 1: return {&launch_solve_t<TW, TH, 1, MODE, HOST>, &occ_solve_t<TW, TH, 1, MODE, HOST>};
        case 2: return {&launch_solve_t<TW, TH, 2, MODE, HOST>, &occ_solve_t<TW, TH, 2, MODE, HOST>};
        case 3: return {&launch_solve_t<TW, TH, 3, MODE, HOST>, &occ_solve_t<TW, TH, 3, MODE, HOST>};
    }
    return {nullptr, nullptr};
}

template <int TW, int TH>
static SolveKernel pick_mode(int kind, int mode, bool host) {
    if (mode == MODE_SYSTOLIC) return host ? pick_kind<TW, TH, MODE_SYSTOLIC, true>(kind)
                                           : pick_kind<TW, TH, MODE_SYSTOLIC, false>(kind);
    if (mode == MODE_FIM) return host ? pick_kind<TW, TH, MODE_FIM, true>(kind)
                                      : pick_kind<TW, TH, MODE_FIM, false>(kind);
    return {nullptr, nullptr};
}

// The 4-waves-per-SIMD form of the 64x64 sweep kernel, for device-scheduled waves wide
// enough to fill four work-groups per CU.
static SolveKernel pick_sweep_wide(int kind) {
    switch (kind) {
        case 0: return {&launch_sweep_t<64, 64, 0, false, 4>, &occ_sweep_t<64, 64, 0, false, 4>};
        case 1: return {&launch_sweep_t<64, 64, 1, false, 4>, &occ_sweep_t<64, 64, 1, false, 4>};
        case 2: return {&launch_sweep_t<64, 64, 2, false, 4>, &occ_sweep_t<64, 64, 2, false, 4>};
        case 3: return {&launch_sweep_t<64, 64, 3, false, 4>, &occ_sweep_t<64, 64, 3, false, 4>};
    }
    return {nullptr, nullptr};
}

SolveKernel get_solve_kernel(int tw, int th, int kind, int mode, bool host, bool wide) {
    if (mode == 2) {
        if (wide && !host && tw == 64 && th == 64) return pick_sweep_wide(kind);
        if (tw == 64 && th == 128) return pick_sweep<64, 128>(kind, host);
        if (tw == 64 && th == 64) return pick_sweep<64, 64>(kind, host);
        if (tw == 64 && th == 32) return pick_sweep<64, 32>(kind, host);
        if (tw == 64 && th == 16) return pick_sweep<64, 16>(kind, host);
        if (tw == 32 && th == 32) return pick_sweep<32, 32>(kind, host);
        return {nullptr, nullptr};
    }
    if (tw == 8 && th == 8) return pick_mode<8, 8>(kind, mode, host);
    if (tw == 16 && th == 16) return pick_mode<16, 16>(kind, mode, host);
    if (tw == 32 && th == 8) return pick_mode<32, 8>(kind, mode, host);
    if (tw == 32 && th == 32) return pick_mode<32, 32>(kind, mode, host);
    return {nullptr, nullptr};
}

// grid: the solving work-groups; with SEL the launch adds kSelCarryGroups carrying ones.
template <int TW, int TH, int KIND, int WPE, bool SEL>
static hipError_t launch_bf_t(const SolveArgs& a, uint32_t wave, uint32_t grid, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1) {
    const uint32_t g = SEL ? grid + kSelCarryGroups : grid;
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve_sweep_bf<TW, TH, KIND, WPE, SEL>), dim3(g), dim3(kSweepThreads),
                              0, s, ev0, ev1, 0, a, wave);
    else
        hipLaunchKernelGGL((k_solve_sweep_bf<TW, TH, KIND, WPE, SEL>), dim3(g), dim3(kSweepThreads), 0,
                           s, a, wave);
    return hipGetLastError();
}

template <int TW, int TH, int KIND, int WPE, bool SEL>
static int occ_bf_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve_sweep_bf<TW, TH, KIND, WPE, SEL>,
                                                     kSweepThreads, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int TW, int TH, int WPE, bool SEL = false>
static BfKernel pick_bf(int kind) {
    switch (kind) {
        case 0: return {&launch_bf_t<TW, TH, 0, WPE, SEL>, &occ_bf_t<TW, TH, 0, WPE, SEL>};
        case 1: return {&launch_bf_t<TW, TH, 1, WPE, SEL>, &occ_bf_t<TW, TH, 1, WPE, SEL>};
        case 2: return {&launch_bf_t<TW, TH, 2, WPE, SEL>, &occ_bf_t<TW, TH, 2, WPE, SEL>};
        case 3: return {&launch_bf_t<TW, TH, 3, WPE, SEL>, &occ_bf_t<TW, TH, 3, WPE, SEL>};
    }
    return {nullptr, nullptr};
}

BfKernel get_bf_kernel(int tw, int th, int kind, int mode, bool wide, bool select) {
    if (mode != 2) return {nullptr, nullptr};
    if (select) {
        // The selection scratch is laid over the tile box: 64 x 64 and 64 x 32 tiles.
        if (tw == 64 && th == 64) return wide ? pick_bf<64, 64, 4, true>(kind) : pick_bf<64, 64, 1, true>(kind);
        if (tw == 64 && th == 32) return pick_bf<64, 32, 1, true>(kind);
        return {nullptr, nullptr};
    }
    if (tw == 64 && th == 64) return wide ? pick_bf<64, 64, 4>(kind) : pick_bf<64, 64, 1>(kind);
    if (tw == 64 && th == 32) return pick_bf<64, 32, 1>(kind);
    if (tw == 64 && th == 16) return pick_bf<64, 16, 1>(kind);
    if (tw == 32 && th == 32) return pick_bf<32, 32, 1>(kind);
    return {nullptr, nullptr};
}

hipError_t launch_wave_end(const SolveArgs& a, uint32_t first, hipStream_t s) {
    hipLaunchKernelGGL(k_wave_end, dim3(1), dim3(1024), 0, s, a, first);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_tile_inst(uint32_t* tinst, const GhcmInst* inst, uint32_t n_inst, hipStream_t s) {
    unsigned blocks = n_inst < 1024 ? n_inst : 1024;
    hipLaunchKernelGGL(k_tile_inst, dim3(blocks), dim3(256), 0, s, tinst, inst, n_inst);
    return hipGetLastError();
}

hipError_t launch_reduce_stats(const uint32_t* tstat, uint32_t tiles, unsigned long long* out,
                              hipStream_t s) {
    unsigned blocks = (tiles + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_reduce_stats, dim3(blocks), dim3(256), 0, s, tstat, tiles, out);
    return hipGetLastError();
}

hipError_t launch_seed_serve(const SolveArgs& a, const uint32_t* seeds, uint32_t n_seeds,
                             uint32_t tw, uint32_t th, hipStream_t s) {
    if (n_seeds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed_serve, dim3((n_seeds + 63) / 64), dim3(64), 0, s, a, seeds, n_seeds, tw, th);
    return hipGetLastError();
}

hipError_t launch_admit_tiles(const SolveArgs& a, const GhcmInst& in, uint32_t slot,
                              uint32_t gen, hipStream_t s, uint32_t tw, uint32_t th, int kind) {
    if (tw && a.edge && th != a.edge_h) return hipErrorInvalidValue;
    const uint32_t count = inst_tiles(in);
    const uint64_t cells = (uint64_t)in.w * in.h * in.d;
    const uint64_t bt = (count + 255) / 256, bc = (cells + 256ull * 16 - 1) / (256ull * 16);
    uint64_t blocks = bt > bc ? bt : bc;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    const dim3 grid((unsigned)blocks), block(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_admit_tiles<0>, grid, block, 0, s, a, in, slot, gen, tw, th); break;
        case 1: hipLaunchKernelGGL(k_admit_tiles<1>, grid, block, 0, s, a, in, slot, gen, tw, th); break;
        case 2: hipLaunchKernelGGL(k_admit_tiles<2>, grid, block, 0, s, a, in, slot, gen, tw, th); break;
        default: hipLaunchKernelGGL(k_admit_tiles<3>, grid, block, 0, s, a, in, slot, gen, tw, th); break;
    }
    return hipGetLastError();
}

hipError_t launch_wake(GhcmCtrl* ctrl, hipStream_t s) {
    hipLaunchKernelGGL(k_wake, dim3(1), dim3(64), 0, s, ctrl);
    return hipGetLastError();
}

hipError_t launch_solve_init(const SolveArgs& a, const GhcmInst* h_inst, uint32_t n_inst,
                             uint64_t max_cells, uint32_t n_select, unsigned long long* totals,
                             hipStream_t s, uint32_t tw, uint32_t th, uint32_t td, int kind) {
    if (tw && a.edge && th * td != a.edge_h) return hipErrorInvalidValue;
    uint64_t bx = (max_cells + 256ull * 16 - 1) / (256ull * 16);
    if (bx < 1) bx = 1;
    if (bx > 1024) bx = 1024;
    const uint32_t by = n_inst < 65535u ? (n_inst ? n_inst : 1u) : 65535u;
    const dim3 grid((unsigned)bx, by), block(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_solve_init<0>, grid, block, 0, s, a, h_inst, n_inst, n_select, totals, tw, th, td); break;
        case 1: hipLaunchKernelGGL(k_solve_init<1>, grid, block, 0, s, a, h_inst, n_inst, n_select, totals, tw, th, td); break;
        case 2: hipLaunchKernelGGL(k_solve_init<2>, grid, block, 0, s, a, h_inst, n_inst, n_select, totals, tw, th, td); break;
        default: hipLaunchKernelGGL(k_solve_init<3>, grid, block, 0, s, a, h_inst, n_inst, n_select, totals, tw, th, td); break;
    }
    return hipGetLastError();
}

hipError_t launch_publish_totals(const unsigned long long* totals, const GhcmCtrl* ctrl,
                                 HostRec* rec, uint32_t tag, hipStream_t s) {
    hipLaunchKernelGGL(k_publish_totals, dim3(1), dim3(64), 0, s, totals, ctrl, rec, tag);
    return hipGetLastError();
}

hipError_t launch_rgb_luminosity(const float* rgb, float* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n / 4 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_rgb_luminosity, dim3((unsigned)blocks), dim3(256), 0, s, rgb, out, n);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    size_t blocks = (bytes + 16 * 256 * 4 - 1) / (16 * 256 * 4);
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, bytes);
    return hipGetLastError();
}

template <int KIND>
static hipError_t launch_solve3d_t(const SolveArgs& a, uint32_t grid, hipStream_t s, hipEvent_t ev0,
                                   hipEvent_t ev1) {
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve3d<KIND>), dim3(grid), dim3(kT3Threads), 0, s, ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL((k_solve3d<KIND>), dim3(grid), dim3(kT3Threads), 0, s, a);
    return hipGetLastError();
}

template <int KIND>
static int occ_solve3d_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve3d<KIND>, kT3Threads, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int KIND>
static hipError_t launch_solve3d_sweep_t(const SolveArgs& a, uint32_t grid, hipStream_t s, hipEvent_t ev0,
                                         hipEvent_t ev1) {
    if (ev0 && ev1)
        hipExtLaunchKernelGGL((k_solve3d_sweep<KIND>), dim3(grid), dim3(kS3Threads), 0, s, ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL((k_solve3d_sweep<KIND>), dim3(grid), dim3(kS3Threads), 0, s, a);
    return hipGetLastError();
}

template <int KIND>
static int occ_solve3d_sweep_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve3d_sweep<KIND>, kS3Threads, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

SolveKernel get_solve3d_kernel(int kind, bool sweep) {
    if (sweep) {
        switch (kind) {
            case 0: return {&launch_solve3d_sweep_t<0>, &occ_solve3d_sweep_t<0>};
            case 1: return {&launch_solve3d_sweep_t<1>, &occ_solve3d_sweep_t<1>};
            case 2: return {&launch_solve3d_sweep_t<2>, &occ_solve3d_sweep_t<2>};
            case 3: return {&launch_solve3d_sweep_t<3>, &occ_solve3d_sweep_t<3>};
            default: return {nullptr, nullptr};
        }
    }
    switch (kind) {
        case 0: return {&launch_solve3d_t<0>, &occ_solve3d_t<0>};
        case 1: return {&launch_solve3d_t<1>, &occ_solve3d_t<1>};
        case 2: return {&launch_solve3d_t<2>, &occ_solve3d_t<2>};
        case 3: return {&launch_solve3d_t<3>, &occ_solve3d_t<3>};
        default: return {nullptr, nullptr};
    }
}

void tile3d_shape(bool sweep, uint32_t& tw, uint32_t& th, uint32_t& td) {
    tw = sweep ? (uint32_t)kS3X : (uint32_t)kT3;
    th = sweep ? (uint32_t)kS3Y : (uint32_t)kT3;
    td = sweep ? (uint32_t)kS3Z : (uint32_t)kT3;
}

hipError_t launch_seed3d(const SolveArgs& a, const uint32_t* seeds, uint32_t n_seeds, uint32_t tw,
                         uint32_t th, uint32_t td, hipStream_t s) {
    if (n_seeds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed3d, dim3((n_seeds + 63) / 64), dim3(64), 0, s, a, seeds, n_seeds, tw, th, td);
    return hipGetLastError();
}

hipError_t launch_seed(const SolveArgs& a, const uint32_t* seeds, uint32_t n_seeds,
                       uint32_t tw, uint32_t th, hipStream_t s) {
    if (n_seeds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed, dim3((n_seeds + 63) / 64), dim3(64), 0, s, a, seeds, n_seeds, tw, th);
    return hipGetLastError();
}

// ---- scratch of the allocation-free solve (ghcm_solve_device_noalloc) ----
// That entry point keeps all of its scheduler state (control block, instance record, tile
// records, wave lists, selection, counters) in this file-scope array instead of buffers the
// library allocates.  Each GPU holds its own copy at its own address, for the life of the
// process; host code never passes the array's name to a kernel, only the address that
// hipGetSymbolAddress returned for the device (noalloc_scratch_init, called with that device
// current when a context is created there).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace ghcm
