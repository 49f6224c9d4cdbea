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
// solves them, carrying the rest to the next list.  Activations append neighbours to
// list[cur^1] once per wave (atomicMax on qwave), sharded over kShards counters.
// Host mode takes an explicit job list and writes one status record per job
// (src/eikonal_hcm.cpp:396-420).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ghcm_dev.h"
#include "kernels.h"

namespace ghcm {

enum { MODE_SYSTOLIC = 0, MODE_FIM = 1 };

template <int KIND>
__device__ __forceinline__ float load_cost(const void* f, size_t i) {
    if constexpr (KIND == 0) {
        return static_cast<const float*>(f)[i];
    } else if constexpr (KIND == 1) {
        return 1.0f / static_cast<const float*>(f)[i];
    } else if constexpr (KIND == 2) {
        return 1.0f / static_cast<float>(static_cast<const _Float16*>(f)[i]);
    } else {
        uint32_t b = static_cast<uint32_t>(static_cast<const uint16_t*>(f)[i]) << 16;
        return 1.0f / __uint_as_float(b);
    }
}

// Godunov upwind update for |grad T| = c with h = 1, cancellation-safe form.
// m = min(a,b), d = |a-b|; if c < d the 1-D update m + c applies (data/eikonal_common.cl:
// 296-297); otherwise T = m + (d + sqrt((sqrt2 c - d)(sqrt2 c + d)))/2, the root of
// (T-a)^2 + (T-b)^2 = c^2 (:299-300) with the discriminant factored and the large
// common part m added once.  Result min'ed with t (:305-308).
__device__ __forceinline__ float godunov(float ax, float ay, float c, float t) {
    float m = fminf(ax, ay);
    float M = fmaxf(ax, ay);
    float r;
    if (__builtin_isinf(M)) {
        r = m + c;
    } else {
        float d = M - m;
        if (c < d) {
            r = m + c;
        } else {
            float s = 1.41421356237309515f * c;
            float q = (s - d) * (s + d);
            r = m + 0.5f * (d + __builtin_sqrtf(q));
        }
    }
    return fminf(t, r);
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Append tile g to list `which` unless it is already there (qwave), in shard g % kShards.
__device__ __forceinline__ void list_append(const SolveArgs& a, uint32_t which, uint32_t g) {
    const uint32_t sh = g % kShards;
    uint32_t slot = atomicAdd(&a.ctrl->count[which][sh * kCounterStride], 1u);
    a.list[((size_t)which * kShards + sh) * a.shard_cap + slot] = g;
}

// Queue tile nb for the next wave unless it is already queued, recording why (border
// side bit) and the smallest candidate time that activated it.
__device__ __forceinline__ void activate(const SolveArgs& a, uint32_t nb, uint32_t wave,
                                         uint32_t cur, uint32_t p, uint32_t mask_bits,
                                         float key) {
    const uint32_t np = p ^ 1u;
    atomicMin(&a.key[(size_t)np * a.tiles + nb], __float_as_uint(key));
    atomicOr(&a.mask[(size_t)np * a.tiles + nb], mask_bits);
    uint32_t old = atomicMax(&a.qwave[nb], wave + 2u);
    if (old < wave + 2u) list_append(a, cur ^ 1u, nb);
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

template <int TW, int TH, int KIND, int MODE, bool HOST>
__device__ void process_tile(const SolveArgs& a, uint32_t g, uint32_t mbits, uint32_t wave,
                             uint32_t cur, uint32_t p, uint32_t job_slot,
                             float* t_s, float* c_s, uint8_t* open_s, uint8_t* add_s) {
    constexpr int PW = TW + 2;
    constexpr int PH = TH + 2;
    constexpr int PP = PW * PH;
    constexpr int P = TW * TH;
    constexpr int CPL = (P + 63) / 64;
    const int lane = threadIdx.x;

    const uint32_t ii = (a.n_inst == 1) ? 0u : a.tinst[g];
    const GhcmInst in = a.inst[ii];
    const uint32_t lt = g - in.tile_base;
    const int tx = (int)(lt % in.tiles_x);
    const int ty = (int)(lt / in.tiles_x);
    const int ox = tx * TW, oy = ty * TH;
    const int W = (int)in.w, H = (int)in.h;
    const int sx = min(TW, W - ox), sy = min(TH, H - oy);

    // ---- stage tile + halo ring into LDS (out of grid -> +INF) ----
    for (int i = lane; i < PP; i += 64) {
        int bx = i % PW, by = i / PW;
        int gx = ox + bx - 1, gy = oy + by - 1;
        float t = __builtin_inff(), c = __builtin_inff();
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            size_t idx = (size_t)gy * (size_t)W + (size_t)gx;
            t = in.time[idx];
            c = load_cost<KIND>(in.field, idx);
        }
        t_s[i] = t;
        c_s[i] = c;
    }
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
        // FIM body (data/eikonal_ghcm_fim.cl:290-372) with a byte open set in LDS.
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
        for (;;) {
            bool any_open = false;
            bool conv[CPL];
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                int i = lane + 64 * k;
                conv[k] = false;
                if (i < P && open_s[i]) {
                    any_open = true;
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
            if (__ballot(any_open) == 0ull) break;
            if (passes >= cap) { unfinished = true; break; }
            ++passes;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                int i = lane + 64 * k;
                if (i < P && conv[k]) {
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
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                int i = lane + 64 * k;
                if (i < P) {
                    uint8_t o = open_s[i];
                    if (conv[k]) o = 0;
                    if (add_s[i]) { o = 1; add_s[i] = 0; }
                    open_s[i] = o;
                }
            }
            __syncthreads();
        }
    }

    __syncthreads();
    // ---- commit improved cells (plain stores: this work-group is the only writer of
    // the tile's cells in this wave), tile min ----
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
                    in.time[(size_t)(oy + y) * (size_t)W + (size_t)(ox + x)] = tn;
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
            a.tstat[g] += 1u;
            a.tstat[(size_t)a.tiles + g] += passes;
            a.tstat[2 * (size_t)a.tiles + g] += improved;
        }
    } else {
        if (lane == 0) {
            a.tmin[g] = tmin;
            a.tstat[g] += 1u;
            a.tstat[(size_t)a.tiles + g] += passes;
            a.tstat[2 * (size_t)a.tiles + g] += improved;
            if (unfinished) activate(a, g, wave, cur, p, kMaskOpenAll, tmin);
        }
        if (lane < 4 && ((side_flags >> lane) & 1u)) {
            const int s = lane;
            int nx = tx + (s == 0 ? -1 : s == 2 ? 1 : 0);
            int ny = ty + (s == 1 ? -1 : s == 3 ? 1 : 0);
            uint32_t nb = in.tile_base + (uint32_t)ny * in.tiles_x + (uint32_t)nx;
            activate(a, nb, wave, cur, p, 1u << (s ^ 2), side_key[s]);
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
    const int lane = threadIdx.x;

    if constexpr (HOST) {
        const uint32_t j = blockIdx.x;
        if (j >= a.n_jobs) return;
        const HostJob job = a.jobs[j];
        process_tile<TW, TH, KIND, MODE, true>(a, job.id, job.border, 0u, 0u, 0u, j,
                                               t_s, c_s, open_s, add_s);
    } else {
        __shared__ uint32_t s_prefix[kShards + 1];
        const GhcmCtrl* ctrl = a.ctrl;
        if (ctrl->done) return;
        const uint32_t c = ctrl->total;
        if (blockIdx.x >= c) return;
        const uint32_t wave = ctrl->wave;
        const uint32_t cur = ctrl->cur;
        const uint32_t thr_lo = ctrl->thr_lo, thr_hi = ctrl->thr_hi, quota = ctrl->quota;
        const uint32_t p = wave & 1u;
        for (int i = lane; i <= (int)kShards; i += 64) s_prefix[i] = ctrl->prefix[i];
        __syncthreads();
        for (uint32_t i = blockIdx.x; i < c; i += gridDim.x) {
            const uint32_t g = list_entry(a, s_prefix, cur, i);
            uint32_t sel = 0, mb = 0;
            if (lane == 0) {
                uint32_t* kp = &a.key[(size_t)p * a.tiles + g];
                uint32_t* mp = &a.mask[(size_t)p * a.tiles + g];
                uint32_t k = *kp;
                sel = (k < thr_lo) ? 1u : 0u;
                if (!sel && k < thr_hi) sel = (atomicAdd(&a.ctrl->quota_used, 1u) < quota) ? 1u : 0u;
                mb = *mp;
                *kp = kInfBits;
                *mp = 0u;
                if (!sel) activate(a, g, wave, cur, p, mb, __uint_as_float(k));  // carry
            }
            sel = __shfl(sel, 0);
            mb = __shfl(mb, 0);
            if (sel) {
                process_tile<TW, TH, KIND, MODE, false>(a, g, mb, wave, cur, p, i,
                                                        t_s, c_s, open_s, add_s);
                __syncthreads();
            }
        }
    }
}

// Closes the previous wave (swap lists), computes the shard prefix of the new list and,
// when it holds more than n tiles, a threshold selecting its n lowest keys: a two-level
// radix histogram over key bits [30:19] then [18:7] (keys are non-negative floats, so
// their bit patterns order like the values).  One work-group of 1024 threads.
__global__ __launch_bounds__(1024) void k_wave_end(SolveArgs a, uint32_t first) {
    __shared__ uint32_t hist[4096];
    __shared__ uint32_t part[1024];
    __shared__ uint32_t s_prefix[kShards + 1];
    __shared__ uint32_t s_bin, s_below;
    GhcmCtrl* ctrl = a.ctrl;
    const int tid = threadIdx.x;
    if (ctrl->done) return;
    uint32_t cur = ctrl->cur;
    const uint32_t wave0 = ctrl->wave;
    if (!first) {
        // list[cur] was consumed by the wave that just ran: empty it and swap.
        if (tid < (int)kShards) ctrl->count[cur][tid * kCounterStride] = 0u;
        cur ^= 1u;
    }
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t sh = 0; sh < kShards; ++sh) {
            s_prefix[sh] = acc;
            acc += ctrl->count[cur][sh * kCounterStride];
        }
        s_prefix[kShards] = acc;
    }
    __syncthreads();
    const uint32_t c = s_prefix[kShards];
    const uint32_t n = ctrl->n_select;
    if (tid <= (int)kShards) ctrl->prefix[tid] = s_prefix[tid];
    if (tid == 0) {
        if (!first) { ctrl->cur = cur; ctrl->wave = wave0 + 1u; }
        ctrl->total = c;
        ctrl->quota_used = 0u;
        if (c == 0u) ctrl->done = 1u;
        else ctrl->busy_waves += 1u;
        if (c <= n) { ctrl->thr_lo = 0xFFFFFFFFu; ctrl->thr_hi = 0xFFFFFFFFu; ctrl->quota = 0u; }
    }
    if (c == 0u || c <= n) return;
    const uint32_t p = (wave0 + (first ? 0u : 1u)) & 1u;
    const uint32_t* key = a.key + (size_t)p * a.tiles;
    uint32_t lo_base = 0u, below = 0u;
    for (int level = 0; level < 2; ++level) {
        const int shift = level == 0 ? 19 : 7;
        for (int i = tid; i < 4096; i += 1024) hist[i] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < c; i += 1024) {
            uint32_t k = key[list_entry(a, s_prefix, cur, i)];
            if (level == 1 && (k >> 19) != (lo_base >> 19)) continue;
            atomicAdd(&hist[(k >> shift) & 0xFFFu], 1u);
        }
        __syncthreads();
        // Each thread owns 4 consecutive bins; exclusive scan of the 1024 partial sums.
        const uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2],
                       h3 = hist[4 * tid + 3];
        part[tid] = h0 + h1 + h2 + h3;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t v = tid >= off ? part[tid - off] : 0u;
            __syncthreads();
            part[tid] += v;
            __syncthreads();
        }
        const uint32_t incl = part[tid] + below;
        const uint32_t excl = incl - (h0 + h1 + h2 + h3);
        if (excl < n && incl >= n) {           // the crossing bin is one of this thread's 4
            uint32_t cum = excl, b = 4 * tid;
            const uint32_t hh[4] = {h0, h1, h2, h3};
            for (int j = 0; j < 4; ++j) {
                if (cum + hh[j] >= n) { b = 4 * tid + j; break; }
                cum += hh[j];
            }
            s_bin = b;
            s_below = cum;
        }
        __syncthreads();
        lo_base |= s_bin << shift;
        below = s_below;
        __syncthreads();
    }
    if (tid == 0) {
        ctrl->thr_lo = lo_base;                 // `below` keys are < lo_base
        ctrl->thr_hi = lo_base + (1u << 7);
        ctrl->quota = n - below;
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
        const uint32_t nt = in.tiles_x * in.tiles_y;
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
    in.time[(size_t)y * in.w + x] = 0.0f;
    const int dx[5] = {0, -1, 1, 0, 0}, dy[5] = {0, 0, 0, -1, 1};
    for (int d = 0; d < 5; ++d) {
        int px = (int)x + dx[d], py = (int)y + dy[d];
        if (px < 0 || py < 0 || px >= (int)in.w || py >= (int)in.h) continue;
        uint32_t t = in.tile_base + (uint32_t)(py / (int)th) * in.tiles_x + (uint32_t)(px / (int)tw);
        atomicMin(&a.key[t], 0u);
        atomicOr(&a.mask[t], kMaskOpenAll);
        uint32_t old = atomicMax(&a.qwave[t], 1u);
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

template <int TW, int TH, int MODE, bool HOST>
static solve_launcher_t pick_kind(int kind) {
    switch (kind) {
        case 0: return &launch_solve_t<TW, TH, 0, MODE, HOST>;
        case 1: return &launch_solve_t<TW, TH, 1, MODE, HOST>;
        case 2: return &launch_solve_t<TW, TH, 2, MODE, HOST>;
        case 3: return &launch_solve_t<TW, TH, 3, MODE, HOST>;
    }
    return nullptr;
}

template <int TW, int TH>
static solve_launcher_t pick_mode(int kind, int mode, bool host) {
    if (mode == MODE_SYSTOLIC) return host ? pick_kind<TW, TH, MODE_SYSTOLIC, true>(kind)
                                           : pick_kind<TW, TH, MODE_SYSTOLIC, false>(kind);
    if (mode == MODE_FIM) return host ? pick_kind<TW, TH, MODE_FIM, true>(kind)
                                      : pick_kind<TW, TH, MODE_FIM, false>(kind);
    return nullptr;
}

solve_launcher_t get_solve_launcher(int tw, int th, int kind, int mode, bool host) {
    if (tw == 8 && th == 8) return pick_mode<8, 8>(kind, mode, host);
    if (tw == 16 && th == 16) return pick_mode<16, 16>(kind, mode, host);
    if (tw == 32 && th == 8) return pick_mode<32, 8>(kind, mode, host);
    if (tw == 32 && th == 32) return pick_mode<32, 32>(kind, mode, host);
    return nullptr;
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

hipError_t launch_seed(const SolveArgs& a, const uint32_t* seeds, uint32_t n_seeds,
                       uint32_t tw, uint32_t th, hipStream_t s) {
    if (n_seeds == 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed, dim3((n_seeds + 63) / 64), dim3(64), 0, s, a, seeds, n_seeds, tw, th);
    return hipGetLastError();
}

}  // namespace ghcm
