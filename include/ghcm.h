/*
 * ghcm.h -- C-ABI of the MI355X-native gHCM eikonal solver (libghcm.so).
 *
 * This layer replaces the reference's OpenCL platform layer (src/util_opencl.h:1-57)
 * and the device half of src/eikonal_hcm.cpp.  Every function is `extern "C"`,
 * takes plain pointers and sizes, never exits the process and never throws; it
 * returns a ghcm_status and leaves a readable message in ghcm_last_error().
 *
 * Field layout (as the reference, src/util_dim.h:19-22 and src/util_field.h):
 * W x H cells, row-major, idx = x + y*W.  A field is either a per-cell COST
 * (the reference's input, src/eikonal.h:40-67; cost = 1/speed with h = 1,
 * src/test_runner.cpp:32-36) or a per-cell SPEED (the reduced-precision storage
 * extension; cost = 1/speed is formed in FP32 on load).  Times are FP32;
 * unreached cells and obstacles (cost = +INF, speed = 0) stay +INF.
 * Seeds are (x, y) pairs of uint32; T(seed) = 0.
 */
#ifndef GHCM_H
#define GHCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GHCM_ABI_VERSION 1

/* Widest grid a solve accepts (cells per row).  Kernels address a tile and its halo as a
 * 32-bit byte offset from the tile's origin, so rows of a tile must span less than 4 GiB. */
#define GHCM_MAX_WIDTH 4194304u

/* Status codes.  Replaces cl::Error + cl_do's exit(-1) (src/util_opencl.h:46-56). */
typedef enum ghcm_status {
    GHCM_OK                  = 0,
    GHCM_ERR_INVALID_ARG     = 1,  /* bad size, null pointer, seed out of grid, bad enum */
    GHCM_ERR_NO_DEVICE       = 2,  /* ordinal out of range / no HIP device              */
    GHCM_ERR_HIP             = 3,  /* a HIP runtime call failed                          */
    GHCM_ERR_OUT_OF_MEMORY   = 4,  /* device or host allocation failed                   */
    GHCM_ERR_NOT_CONVERGED   = 5,  /* wave cap reached with tiles still queued           */
    GHCM_ERR_UNSUPPORTED     = 6,  /* valid enum, but no kernel built for that combination */
    GHCM_ERR_OVER_BUDGET     = 7,  /* serving: request larger than the HBM budget         */
    GHCM_ERR_NOT_FOUND       = 8,  /* serving: unknown ticket                             */
    GHCM_ERR_STOPPED         = 9,  /* serving: server stopped                             */
    GHCM_ERR_TIMEOUT         = 10  /* serving: wait timed out, the request is still running */
} ghcm_status;

/* In-tile solver style (HcmConfig::Kernel, src/eikonal.h:13). */
typedef enum ghcm_kernel {
    GHCM_KERNEL_SYSTOLIC = 0,      /* data/eikonal_ghcm_systolic.cl */
    GHCM_KERNEL_FIM      = 1,      /* data/eikonal_ghcm_fim.cl      */
    GHCM_KERNEL_SWEEP    = 2       /* systolic relaxation by four concurrent fast-sweeping
                                      orders, one wave64 each; tiles 64x64, 64x32, 64x16,
                                      32x32 */
} ghcm_kernel;

/* Who picks the n lowest-min-time tiles of a wave. */
typedef enum ghcm_scheduler {
    GHCM_SCHED_DEVICE_TOPN = 0,    /* device-resident wave loop with device top-n      */
    GHCM_SCHED_HOST_PQ     = 1     /* host std::priority_queue, src/eikonal_hcm.cpp:343-482 */
} ghcm_scheduler;

/* HcmConfig::CacheMode (src/eikonal.h:14).  Accepted for source compatibility:
 * tiles are always staged in LDS with their halo, so every mode behaves as BOTH. */
typedef enum ghcm_cache_mode {
    GHCM_CACHE_NONE = 0, GHCM_CACHE_TIME = 1, GHCM_CACHE_COST = 2, GHCM_CACHE_BOTH = 3
} ghcm_cache_mode;

/* Storage of the input field. */
typedef enum ghcm_field_kind {
    GHCM_FIELD_COST_F32   = 0,     /* reference input: FP32 cost                  */
    GHCM_FIELD_SPEED_F32  = 1,     /* FP32 speed, cost = 1/speed in FP32          */
    GHCM_FIELD_SPEED_F16  = 2,     /* IEEE half speed                             */
    GHCM_FIELD_SPEED_BF16 = 3      /* bfloat16 speed                              */
} ghcm_field_kind;

/* Solver configuration.  Mirrors HcmConfig (src/eikonal.h:12-27) plus the
 * MI355X-specific knobs.  Fill with ghcm_config_default() first. */
typedef struct ghcm_config {
    uint32_t tile_w, tile_h;        /* tile shape; block_size x block_size in the reference.
                                       Built shapes: 8x8, 16x16, 32x8, 32x32          */
    float    workers_overload;      /* reference knob (WG = overload x 64); one wave64
                                       solves a tile here, so it is recorded only     */
    float    compute_unit_overload; /* n = max(1, overload x #CUs) tiles per wave     */
    uint32_t n_active;              /* if non-zero, overrides n                       */
    float    min_change;            /* fim_minimum_change (epsilon), >= 0             */
    int32_t  cache_mode;            /* ghcm_cache_mode (no effect, see above)         */
    int32_t  kernel;                /* ghcm_kernel                                    */
    int32_t  scheduler;             /* ghcm_scheduler                                 */
    uint32_t inner_iter_cap;        /* in-tile passes per visit, 0 = until converged  */
    uint32_t max_waves;             /* safety cap on waves, 0 = automatic             */
    uint32_t waves_per_poll;        /* device scheduler: waves kept enqueued ahead of
                                       the last executed wave boundary (at least 2),
                                       0 = 4                                          */
    uint32_t profile_kernels;       /* 1: time every tile-solve launch with HIP events
                                       (fills solve_kernel_us / solve_launches)       */
    uint32_t launch_mode;           /* device scheduler: 0 = one tile-solve launch per
                                       wave (default), 1 = one persistent cooperative
                                       launch runs every wave where built (sweep kernel;
                                       its grid barrier pays agent-scope fences),
                                       2 = chunks of waves replayed as a HIP graph (host
                                       scheduler: each wave's job upload, launch and
                                       status read-back replayed as one graph),
                                       3 = one launch per wave: the tile-solve kernel's
                                       last work-group runs the wave boundary (sweep),
                                       4 = boundary-free waves (sweep kernel): one
                                       tile-solve launch per wave and no wave-boundary
                                       kernel; every queued tile is solved in its wave
                                       (no top-n cut), n sets the launch's work-groups,
                                       5 = boundary-free waves with the top-n cut made by
                                       the tile-solve launch itself (the wave boundary's
                                       selection rule; 64x64 and 64x32 sweep tiles, other
                                       shapes run as 0); the allocation-free entry point
                                       takes 0 and 4 only                               */
} ghcm_config;

/* Solve statistics (EikonalResult durations, src/eikonal.h:31-36, plus counters). */
typedef struct ghcm_stats {
    double   duration_total_us;     /* wall time of the solve call                    */
    double   duration_device_us;    /* GPU time between the first and last event      */
    uint64_t waves;                 /* waves that processed at least one tile         */
    uint64_t tile_visits;           /* tiles solved (sum over waves)                  */
    uint64_t tiles_total;           /* tiles in the grid(s)                           */
    uint64_t cells;                 /* W*H summed over instances                      */
    uint64_t inner_iterations;      /* in-tile passes summed over visits              */
    uint64_t cells_improved;        /* cell commits that lowered a time               */
    uint32_t converged;             /* 1 if the queue drained                         */
    uint32_t tile_w, tile_h;        /* tile shape used                                */
    uint32_t n_active;              /* n used                                         */
    uint32_t bytes_per_visit;       /* algorithmic HBM bytes of one tile visit        */
    uint32_t tile_d;                /* tile depth (3-D solves; 1 otherwise)           */
    double   solve_kernel_us;       /* profile_kernels: summed tile-solve launch time */
    uint64_t solve_launches;        /* profile_kernels: tile-solve launches timed     */
} ghcm_stats;

/* Device description (OclDevInfo fields, src/util_opencl.h:4-16). */
typedef struct ghcm_device_info {
    char     name[256];
    char     arch[64];              /* e.g. "gfx950"                                  */
    int32_t  ordinal;
    int32_t  compute_units;
    int32_t  clock_khz;
    int32_t  max_workgroup;
    uint64_t lds_bytes;             /* per workgroup                                  */
    uint64_t global_mem_bytes;
    int32_t  pci_bus_id;
    int32_t  reserved;
} ghcm_device_info;

typedef struct ghcm_ctx ghcm_ctx;   /* owns a HIP stream and scratch; one host thread at a time */

/* ---- library / device ---------------------------------------------------- */
int         ghcm_abi_version(void);
const char* ghcm_status_string(int status);
const char* ghcm_last_error(void);  /* thread-local message of the last failure   */
void        ghcm_config_default(ghcm_config* cfg);
int         ghcm_device_count(int* count);
int         ghcm_device_info_get(int ordinal, ghcm_device_info* out);

/* ---- context (HcmSetup, src/eikonal.h:53-63) ------------------------------ */
int ghcm_ctx_create(int ordinal, const ghcm_config* cfg, ghcm_ctx** out);
int ghcm_ctx_destroy(ghcm_ctx* ctx);
int ghcm_ctx_get_config(const ghcm_ctx* ctx, ghcm_config* out);
int ghcm_ctx_set_config(ghcm_ctx* ctx, const ghcm_config* cfg);

/* ---- plain solve (eikonal_hcm_opencl_ex, src/eikonal_hcm.cpp:186-529) ------
 * Host buffers: `field` (W*H elements of the given kind) is copied in, the
 * result is copied out into `time_out` (W*H floats).  Buffers are borrowed
 * for the duration of the call only.  `stats` may be NULL. */
int ghcm_solve(ghcm_ctx* ctx, const void* field, int field_kind, uint32_t w, uint32_t h,
               const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
               ghcm_stats* stats);

/* Device buffers: `d_field` and `d_time` are device pointers on the context's
 * device (e.g. from hipMalloc or a torch tensor).  Runs on the context stream
 * and returns when the solve is complete.  Seeds are host memory. */
int ghcm_solve_device(ghcm_ctx* ctx, const void* d_field, int field_kind,
                      uint32_t w, uint32_t h, const uint32_t* seeds_xy, uint32_t n_seeds,
                      float* d_time, ghcm_stats* stats);

/* ---- allocation-free solve ----------------------------------------------------
 * No counterpart in the reference: ghcm_solve_device for callers that cannot allocate
 * (tight serving loops, code that may not call hipMalloc, harnesses that forbid it).  Same
 * arguments and results; it calls no hipMalloc, hipHostMalloc, hipFree or other allocating
 * routine.  Its scheduler state lives in a static device array of 16 MiB per GPU, held
 * until the process ends, whose address ghcm_ctx_create looks up once per device; the
 * pinned records it uses are reserved by ghcm_ctx_create too.  Limits: the grid's tiles must
 * fit the array (52 B per tile: 322,602 tiles, a 4096^2 grid in 8x8 tiles; the sweep kernel
 * adds 16 B per tile row for its edge columns: 15,590 tiles of 64x64, a 7,900^2 grid; more is
 * refused with GHCM_ERR_INVALID_ARG), at most 4,096 seeds, the device
 * scheduler with launch_mode 0 or 4 and profile_kernels off (else GHCM_ERR_UNSUPPORTED).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU, from any
 * context, race; each call starts by clearing what must start at zero. */
int ghcm_solve_device_noalloc(ghcm_ctx* ctx, const void* d_field, int field_kind,
                              uint32_t w, uint32_t h, const uint32_t* seeds_xy,
                              uint32_t n_seeds, float* d_time, ghcm_stats* stats);

/* ---- three-dimensional solve (the K = 3 generalisation of the update,
 * eikonal_solve<K>, src/eikonal.h:73-118 and data/eikonal_common.cl:221-285; the
 * reference declares K <= 3 but solves only 2-D grids) ----------------------------
 * `field` and the time are W*H*D cells, idx = x + W*(y + H*z); `seeds_xyz` holds
 * n_seeds (x, y, z) triples (host memory).  kernel GHCM_KERNEL_SWEEP solves 64x8x8
 * tiles with the eight octant sweeps (inner_iter_cap counts rounds of all eight), the
 * other kernels 8x8x8 tiles with in-tile passes; the device scheduler runs the wave loop:
 * n_active, min_change, inner_iter_cap, waves_per_poll, max_waves and launch_mode 2
 * (graph replay) apply, the tile shape fields do not (stats report tile_w/h/d), and
 * the host scheduler returns GHCM_ERR_UNSUPPORTED.  Host-buffer and device-buffer forms
 * as for the 2-D solve. */
int ghcm_solve3d(ghcm_ctx* ctx, const void* field, int field_kind, uint32_t w, uint32_t h,
                 uint32_t d, const uint32_t* seeds_xyz, uint32_t n_seeds, float* time_out,
                 ghcm_stats* stats);
int ghcm_solve3d_device(ghcm_ctx* ctx, const void* d_field, int field_kind, uint32_t w,
                        uint32_t h, uint32_t d, const uint32_t* seeds_xyz, uint32_t n_seeds,
                        float* d_time, ghcm_stats* stats);

/* ---- device memory on the context's device (plumbing for ghcm_solve_device and
 * ghcm_solve_batch, so hosts need no other GPU runtime in-process) ---------------- */
int ghcm_dev_alloc(ghcm_ctx* ctx, size_t bytes, void** d_out);
int ghcm_dev_free(ghcm_ctx* ctx, void* d_ptr);
int ghcm_dev_upload(ghcm_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int ghcm_dev_download(ghcm_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ghcm_dev_copy(ghcm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---- batched solve: many uneven instances in one device wave loop ---------- */
typedef struct ghcm_instance {
    const void*     d_field;        /* device pointer, field_kind of the batch        */
    float*          d_time;         /* device pointer, W*H floats                     */
    uint32_t        w, h;
    const uint32_t* seeds_xy;       /* host pointer                                   */
    uint32_t        n_seeds;
    uint32_t        reserved;
} ghcm_instance;

int ghcm_solve_batch(ghcm_ctx* ctx, int field_kind, const ghcm_instance* inst,
                     uint32_t n_inst, ghcm_stats* stats);

/* ---- autotuner --------------------------------------------------------------
 * Picks the tile shape/kernel, n, in-tile round cap and launch mode for a workload by
 * timing candidate configurations on it (device time, best of `reps`).  The reference
 * sweeps the same kind of grid by hand (src/test_runner.cpp:220-285). */
typedef struct ghcm_tune_candidate {
    uint32_t tile_w, tile_h;
    int32_t  kernel;                /* ghcm_kernel                                    */
    uint32_t n_active;
    uint32_t inner_iter_cap;
    uint32_t launch_mode;
    double   device_us;             /* out: best device time (INFINITY if it failed)  */
    uint64_t tile_visits;           /* out                                            */
    uint32_t converged;             /* out                                            */
    int32_t  status;                /* out: ghcm_status of the last run               */
} ghcm_tune_candidate;

/* Fills `out` (capacity `cap`) with the default search space for `compute_units` CUs:
 * systolic and FIM tiles 8x8/16x16/32x8, sweep tiles 64x64/64x32/32x32, n in
 * {1,2,4,16,64} x CUs, round caps and launch modes.  *n_out = number written. */
int ghcm_tune_default_space(int compute_units, ghcm_tune_candidate* out, uint32_t cap,
                            uint32_t* n_out);

/* Times every candidate on the instances (device pointers, as ghcm_solve_batch) and
 * leaves the fastest converged one as the context configuration (*best = its index). */
int ghcm_autotune(ghcm_ctx* ctx, int field_kind, const ghcm_instance* inst, uint32_t n_inst,
                  ghcm_tune_candidate* cands, uint32_t n_cands, uint32_t reps, uint32_t* best);

/* ---- serving: continuous batching -------------------------------------------------
 * No counterpart in the reference, which solves one field per call
 * (src/eikonal_hcm.cpp:186-529); this is the serving form of the same wave loop.
 * A server runs one device wave loop that never restarts between requests.  A request
 * is admitted at a wave boundary when an instance slot, its tile state and its field +
 * time buffer inside the HBM arena (`hbm_budget_bytes`) are free, first come first
 * served; its tiles then join the same top-n selection as every live request's.  It
 * completes at the first wave boundary at which none of its tiles is queued, and its
 * times are copied into `time_out`.
 * Threading: submit / wait / get_stats may be called from any threads; stop must not
 * race with them. */
typedef struct ghcm_serve_config {
    uint64_t hbm_budget_bytes;      /* arena for request fields and times, 0 = 8 GiB  */
    uint32_t max_tiles;             /* tile-state capacity, 0 = from the budget       */
    uint32_t max_slots;             /* requests on the device at once, 0 = 256, <= 1024 */
    int32_t  field_kind;            /* ghcm_field_kind of every request               */
    uint32_t reserved[3];
} ghcm_serve_config;

typedef struct ghcm_serve_result {
    int32_t  status;                /* ghcm_status of the request                     */
    uint32_t w, h;
    uint32_t reserved;
    uint64_t waves;                 /* loop waves from admission to completion
                                       (counted in whole chunks of waves_per_poll)    */
    double   queue_us;              /* submit -> admission                            */
    double   latency_us;            /* submit -> times written to time_out            */
} ghcm_serve_result;

typedef struct ghcm_serve_stats {
    uint64_t submitted, admitted, completed, pending;
    uint64_t waves_enqueued;        /* loop waves launched                            */
    uint64_t busy_waves;            /* waves that had at least one queued tile        */
    uint64_t loop_wave;             /* the loop's wave counter at the last poll       */
    uint64_t peak_live;             /* most requests on the device at once            */
    uint64_t budget_bytes, bytes_in_use, peak_bytes;
    uint64_t max_tiles;
    uint32_t slots, n_active, waves_per_chunk;
    uint32_t waiters;               /* ghcm_serve_wait calls in progress              */
} ghcm_serve_stats;

typedef struct ghcm_server ghcm_server;

void ghcm_serve_config_default(ghcm_serve_config* scfg);
/* cfg: solver configuration (device top-n scheduler; launch_mode 2 replays each chunk
 * of waves as one HIP graph, 1 behaves as 0).  Allocates the arena and tile state. */
int ghcm_serve_start(int ordinal, const ghcm_config* cfg, const ghcm_serve_config* scfg,
                     ghcm_server** out);
/* Queues one solve.  `field` (W*H elements of the server's field kind) and the seeds
 * are copied before this returns (the field into pinned staging); `time_out` (W*H
 * floats, host memory) receives the times when the ticket is collected by
 * ghcm_serve_wait.  GHCM_ERR_OVER_BUDGET if the request can never fit. */
int ghcm_serve_submit(ghcm_server* s, const void* field, uint32_t w, uint32_t h,
                      const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
                      uint64_t* ticket);
/* Device-resident request: `d_field` and `d_time` are device buffers on the server's
 * device; the times are written into `d_time` in place and no arena bytes are used.
 * Both buffers must stay valid until the request completes. */
int ghcm_serve_submit_device(ghcm_server* s, const void* d_field, uint32_t w, uint32_t h,
                             const uint32_t* seeds_xy, uint32_t n_seeds, float* d_time,
                             uint64_t* ticket);
/* Waits up to timeout_s seconds (< 0: no limit).  Returns the request's status (and
 * collects the ticket) or GHCM_ERR_TIMEOUT (the ticket stays valid).  One waiter per
 * ticket: a second one gets GHCM_ERR_INVALID_ARG, an unknown or collected ticket
 * GHCM_ERR_NOT_FOUND.  `out` is written only when this call collects the ticket
 * (out->status >= 0), and only then may the caller reuse the request's `time_out`:
 * a host request's times are copied into it by the collecting call. */
int ghcm_serve_wait(ghcm_server* s, uint64_t ticket, double timeout_s, ghcm_serve_result* out);
int ghcm_serve_get_stats(ghcm_server* s, ghcm_serve_stats* out);
/* drain != 0: every submitted request completes first; 0: requests not yet complete
 * end with GHCM_ERR_STOPPED.  Waits blocked in ghcm_serve_wait return first; then the
 * server is freed (no call may use it after this one starts returning). */
int ghcm_serve_stop(ghcm_server* s, int drain);

/* ---- synthetic fields (seeded, documented generators) ---------------------- */
/* splitmix64-based, defined in ghcm_amd/csrc/field_gen.cpp.  Output is FP64 speed
 * (H x W, row-major) so every storage precision derives from one field.
 * ring:   concentric ring walls of speed `wall_speed` (one gap each) on speed 1, the
 *         256^2 geometry scaled by `scale`;
 * smooth: standard-normal noise smoothed by a Gaussian of `sigma` cells (reflecting
 *         boundary), mapped log-linearly onto [speed_lo, speed_hi];
 * points: n uniform cells.  */
int ghcm_gen_ring_speed(uint32_t w, uint32_t h, uint64_t seed, double wall_speed,
                        double scale, double* speed_out);
int ghcm_gen_smooth_speed(uint32_t w, uint32_t h, uint64_t seed, double sigma,
                          double speed_lo, double speed_hi, double* speed_out);
int ghcm_gen_points(uint32_t w, uint32_t h, uint64_t seed, uint32_t n, uint32_t* xy_out);

/* ---- test support: the scheduler's bookkeeping, callable on its own -----------
 * Neither function has a counterpart in the reference; they expose the library's own
 * tile bookkeeping (the block grid and border masks of src/eikonal_hcm.cpp:203-224 and
 * :442-469, and the wave pop of :344-358) so tests can pin it directly. */

/* Geometry of global tile `tile` of a w x h instance in tile_w x tile_h tiles whose ids
 * start at tile_base: row-major ids, partial edge tiles, the neighbour across each side
 * (0 x-, 1 y-, 2 x+, 3 y+; 0xFFFFFFFF at the border) and the border bit that neighbour
 * receives when this tile's edge on that side changes.  Host only (no device needed). */
typedef struct ghcm_tile_desc {
    uint32_t tiles_x, tiles_y;      /* tiles of the instance                          */
    uint32_t tx, ty;                /* this tile's coordinates                        */
    uint32_t x0, y0, w, h;          /* its cells: origin and extent                   */
    uint32_t neighbour[4];          /* global id across side s, or 0xFFFFFFFF         */
    uint32_t border_bit[4];         /* bit set on neighbour[s] by a change on side s  */
    uint32_t seed_tile;             /* tile holding cell (x0 + w - 1, y0 + h - 1)     */
} ghcm_tile_desc;
int ghcm_tile_info(uint32_t w, uint32_t h, uint32_t tile_w, uint32_t tile_h, uint32_t tile_base,
                   uint32_t tile, ghcm_tile_desc* out);

/* One device wave boundary (the top-n selection of k_wave_end) run on a given list:
 * `count` distinct tiles of a `tiles_total`-tile grid with their keys (the smallest
 * candidate time that queued them, >= 0 or +INF) and border masks, queued for wave 0 in
 * list order over the 32 shards (tile g goes to shard g % 32).  n: tiles per wave.
 * handoff = 0 runs the solve's first boundary (unselected tiles are carried to the next
 * list by the boundary itself); handoff = 1 runs a later boundary, which leaves the
 * unselected tiles of a long list (more than 2,048 entries) to the next tile launch.
 * Outputs (capacity `count` each):
 *   jobs[0, *n_jobs)            the selected tiles, in list order;
 *   handed[0, *n_handed)        tiles left to the tile launch (handoff only), list order;
 *   next[0, *n_next)            the next wave's list as the boundary left it, shard by
 *                               shard, with next_keys / next_masks of each entry.
 * A test entry point: it overwrites the context's scheduler buffers. */
int ghcm_debug_select(ghcm_ctx* ctx, uint32_t tiles_total, const uint32_t* tiles,
                      const float* keys, const uint32_t* masks, uint32_t count, uint32_t n,
                      int handoff, uint32_t* jobs, uint32_t* n_jobs, uint32_t* handed,
                      uint32_t* n_handed, uint32_t* next, float* next_keys,
                      uint32_t* next_masks, uint32_t* n_next);

#ifdef __cplusplus
}
#endif
#endif
