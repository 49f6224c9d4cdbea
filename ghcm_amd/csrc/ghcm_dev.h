// ghcm_dev.h -- structures shared by the host scheduler and the gfx950 kernels.
//
// HBM layout of one solve (or one batch of instances):
//   field[W*H]       per instance, COST_F32 / SPEED_{F32,F16,BF16}, row-major
//   time [W*H]       per instance, FP32, row-major
//   tile arrays, one entry per tile of every instance (global tile id g):
//     qwave[g]       u32  wave+1 for which g is queued (0 = never); monotone, so a
//                         tile is appended at most once per wave list (atomicMax)
//     key[2][g]      u32  float bits of the smallest candidate time that activated
//                         g; double-buffered by wave parity
//     mask[2][g]     u32  border bits (1 << side) + OPEN_ALL; double-buffered
//     tmin[g]        f32  smallest own-cell time after the last visit
//     tinst[g]       u32  instance index (batches only)
//     tstat[3][g]    u32  visits / in-tile passes / improved cells of tile g
//                         (owner-only increments; summed once per solve)
//   list[3][kShards][shard_cap]  u32  wave lists, sharded by g % kShards so list
//                         appends spread over kShards counters (each on its own
//                         128-byte line) instead of contending on one address.
//                         With a wave-boundary kernel, list[cur] is read by wave w and
//                         list[cur^1] collects the activations for wave w+1 (two lists).
//                         Boundary-free waves rotate three: wave w reads list[w % 3],
//                         appends to list[(w+1) % 3] and empties list[(w+2) % 3].
//   ctrl             GhcmCtrl, one per solve
#pragma once

#include <stdint.h>

namespace ghcm {

constexpr uint32_t kInfBits   = 0x7F800000u;  // +INF as uint32 bits
constexpr uint32_t kMaskOpenAll = 1u << 4;    // open every cell on the next visit
// Side numbering of the reference (src/eikonal_hcm.cpp:18, data/eikonal_common.cl:361):
// 0: x-, 1: y-, 2: x+, 3: y+ ; opposite(s) = s ^ 2.

// One instance: field and time of W x H x D cells, idx = x + W (y + H z); a 2-D instance has
// D = 1 and tiles_z = 1.  Tile (tx, ty, tz) has the global id tile_base + tx + tiles_x (ty +
// tiles_y tz).
struct GhcmInst {
    const void* field;
    float*      time;
    uint32_t    w, h;
    uint32_t    tiles_x, tiles_y;
    uint32_t    tile_base;
    uint32_t    d;
    uint32_t    tiles_z;
    uint32_t    pad;
};
static_assert(sizeof(GhcmInst) == 48, "instance record layout");

__host__ __device__ inline uint32_t inst_tiles(const GhcmInst& in) {
    return in.tiles_x * in.tiles_y * in.tiles_z;
}

constexpr uint32_t kShards = 32;        // list shards
constexpr uint32_t kCounterStride = 32; // u32 between shard counters (128 B)

constexpr uint32_t kMaxServeSlots = 1024; // instance slots of a serving loop
constexpr uint32_t kBarGroup = 32;      // work-groups per arrival counter (persistent loop)
constexpr uint32_t kMaxBarGroups = 64;

struct GhcmCtrl {
    uint32_t wave;          // index of the wave being (or about to be) processed
    uint32_t cur;           // list[cur] holds the tiles queued for `wave`
    uint32_t done;          // 1 once a wave starts with an empty list
    uint32_t n_select;      // n
    uint32_t n_jobs;        // tiles selected for `wave`
    uint32_t jobs_from_list;// 1: the jobs are list[cur] itself; 0: jobs[] (compacted top-n)
    uint32_t busy_waves;    // waves that had work
    uint32_t total;         // entries in list[cur]
    uint32_t wave_ready;    // persistent loop: wave index whose selection is published
    uint32_t abort;         // persistent loop: a barrier wait timed out
    uint32_t carried;       // tiles carried over by the top-n selection (all waves)
    uint32_t ticks;         // k_wave_end launches executed, the idle ones included
    uint32_t n_carry;       // tiles a long selection left to the wave's tile launch to carry
                            // (jobs[n_jobs, n_jobs + n_carry))
    uint32_t pad0[3];
    uint32_t prefix[kShards + 1];   // exclusive prefix of the shard counts of list[cur]
    uint32_t pad1[31];
    uint32_t count[3][kShards * kCounterStride];
    uint32_t arrive_top;
    uint32_t pad2[31];
    uint32_t arrive[kMaxBarGroups * kCounterStride];
};

// Host-scheduled wave: job and status records (src/eikonal_hcm.cpp:22-32,
// data/eikonal_common.cl:34-44).  Status carries the smallest time computed by
// the visit (own cells and border solves) and one dirty flag per side.
struct HostJob {
    uint32_t id;
    uint32_t border;        // 1 << side, plus kMaskOpenAll for seed tiles
};

struct HostStatus {
    uint32_t solves;        // in-tile passes
    uint32_t cells_updated;
    float    smallest_time;
    uint32_t border[4];
    uint32_t requeue;       // 1 if the visit hit the pass cap before converging
};

// Progress record in mapped, coherent pinned host memory.  k_wave_end writes it with
// system-scope stores on every launch; the host polls it instead of copying the control
// block back or waiting on events, so the wave loop has no runtime copy in it.
struct HostRec {
    // (done << 32) | ticks: k_wave_end launches executed, and whether the loop found its
    // list empty; one 64-bit store per launch
    unsigned long long progress;
    uint32_t wave;                  // refreshed every 16 ticks, when done is set and by
    uint32_t busy_waves;            // k_publish_totals
    uint32_t totals_tag;            // set by k_publish_totals once totals[] is written
    uint32_t pad[11];
    unsigned long long totals[3];   // tile visits, in-tile passes, cells improved
    unsigned long long pad2[5];
    // serving: per slot (generation << 32) | tiles queued at the last wave boundary; a
    // generation is reported only by boundaries after the slot's admission
    unsigned long long slot[kMaxServeSlots];
};

struct SolveArgs {
    const GhcmInst* inst;
    uint32_t        n_inst;
    const uint32_t* tinst;
    uint32_t*       qwave;
    uint32_t*       key;     // [2][tiles]
    uint32_t*       mask;    // [2][tiles]
    float*          tmin;
    uint32_t*       list;    // [3][kShards][shard_cap] (the first two with a boundary kernel)
    uint32_t        tiles;
    uint32_t        shard_cap;
    uint32_t*       tstat;   // [3][tiles]
    uint32_t*       jobs;    // compacted top-n selection followed by the tiles a long selection
                             // hands to the tile launch: capacity `tiles` words
    HostRec*        rec;     // device view of the progress record (null: not written)
    const uint32_t* sgen;    // serving: admission generation per slot
    uint32_t        n_slots; // serving: instance slots (<= kMaxServeSlots), 0 otherwise
    GhcmCtrl*       ctrl;
    float           eps;
    uint32_t        iter_cap;
    // host-scheduled mode
    const HostJob*  host_jobs;
    HostStatus*     status;
    uint32_t        n_host_jobs;
    // Edge columns of every tile (sweep kernel; null otherwise), TH floats per column:
    //   edge[g][0][r]  time of tile g's first column (x = ox) at row oy + r,
    //   edge[g][1][r]  time of its last column (x = ox + TW - 1), full-width tiles only;
    //   cedge[g][s][r] the cost of the same cells, filled once per solve.
    // A tile's left and right halo columns are its neighbours' edge columns: staging reads
    // them as two contiguous runs instead of one cache line per row.  The time edges are
    // written by the tile's own commit, seeds and initialisation, so they always equal the
    // time field.
    float*          edge;
    const float*    cedge;
    uint32_t        edge_h;  // TH of the edge columns (0: no edges)
};

}  // namespace ghcm
