// kernels.h -- host-side entry points of kernels.hip (launch wrappers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ghcm_dev.h"

namespace ghcm {

// ev0/ev1: optional events stamped at the start and end of the dispatch.
using solve_launcher_t = hipError_t (*)(const SolveArgs&, uint32_t grid, hipStream_t,
                                        hipEvent_t ev0, hipEvent_t ev1);

struct SolveKernel {
    solve_launcher_t launch;          // nullptr when no kernel is built for the combination
    int (*blocks_per_cu)();           // resident work-groups per CU (occupancy)
};

// mode: 0 systolic, 1 FIM, 2 sweep.  host: job list given by the host scheduler.  wide: the
// waves hold many more tiles than CUs (selects a higher-occupancy form where one is built).
SolveKernel get_solve_kernel(int tile_w, int tile_h, int kind, int mode, bool host,
                             bool wide = false);
// Three-dimensional tile solve (device scheduler) for a field kind: sweep, 64 x 8 x 8 tiles
// relaxed by the eight octant sweeps; otherwise 8 x 8 x 8 tiles relaxed by in-tile passes.
SolveKernel get_solve3d_kernel(int kind, bool sweep);
// Tile extents (x, y, z) of the 3-D kernels.
void tile3d_shape(bool sweep, uint32_t& tw, uint32_t& th, uint32_t& td);
// Tile solve with the wave boundary folded in (one launch per wave); sweep kernel only.
SolveKernel get_fused_kernel(int tile_w, int tile_h, int kind, int mode);

// Boundary-free waves (sweep kernel): one tile launch per wave, the wave index as argument; no
// wave-boundary kernel.  Without `select` every queued tile is solved in its wave (no top-n
// cut); with it the launch makes the top-n cut itself (three activation records; 64 x 64 and
// 64 x 32 tiles), and `grid` counts the solving work-groups (the launch adds the carrying ones).
struct BfKernel {
    hipError_t (*launch)(const SolveArgs&, uint32_t wave, uint32_t grid, hipStream_t,
                         hipEvent_t ev0, hipEvent_t ev1);
    int (*blocks_per_cu)();
};
BfKernel get_bf_kernel(int tile_w, int tile_h, int kind, int mode, bool wide, bool select = false);

// Persistent device wave loop (cooperative launch); null when not built for the combination.
struct PersistentKernel {
    hipError_t (*launch)(const SolveArgs&, uint32_t grid, hipStream_t);
    int (*blocks_per_cu)();
};
PersistentKernel get_persistent_kernel(int tile_w, int tile_h, int kind, int mode);

hipError_t launch_wave_end(const SolveArgs& a, uint32_t first, hipStream_t s);

// Scratch of the allocation-free solve: a static device array of kNoAllocScratchBytes per
// GPU.  noalloc_scratch_init (device current) looks up its address on that device once;
// noalloc_scratch returns it (null before the lookup).
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
hipError_t launch_tile_inst(uint32_t* tinst, const GhcmInst* inst, uint32_t n_inst, hipStream_t s);
hipError_t launch_reduce_stats(const uint32_t* tstat, uint32_t tiles, unsigned long long* out,
                              hipStream_t s);
hipError_t launch_seed_serve(const SolveArgs& a, const uint32_t* seeds_xyi, uint32_t n_seeds,
                             uint32_t tile_w, uint32_t tile_h, hipStream_t s);
// One admitted request: its slot entry, tile states and tile map, its time field at +INF and,
// with tw != 0 (th rows per edge column, cost kind `kind`), its tiles' edge columns.
hipError_t launch_admit_tiles(const SolveArgs& a, const GhcmInst& in, uint32_t slot,
                              uint32_t gen, hipStream_t s, uint32_t tw = 0, uint32_t th = 0,
                              int kind = 0);
hipError_t launch_wake(GhcmCtrl* ctrl, hipStream_t s);
// h_inst: device view of the mapped host instance table.  tw != 0: also the edge columns of
// every instance's tiles (tw x th x td tiles, cost kind `kind`; nothing when a.edge is null).
hipError_t launch_solve_init(const SolveArgs& a, const GhcmInst* h_inst, uint32_t n_inst,
                             uint64_t max_cells, uint32_t n_select, unsigned long long* totals,
                             hipStream_t s, uint32_t tw = 0, uint32_t th = 0, uint32_t td = 1,
                             int kind = 0);
hipError_t launch_publish_totals(const unsigned long long* totals, const GhcmCtrl* ctrl,
                                 HostRec* rec, uint32_t tag, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// n pixels of interleaved RGB floats -> n luminosity floats.
hipError_t launch_rgb_luminosity(const float* rgb, float* out, size_t n, hipStream_t s);
hipError_t launch_seed(const SolveArgs& a, const uint32_t* seeds_xyi, uint32_t n_seeds,
                       uint32_t tile_w, uint32_t tile_h, hipStream_t s);
// Seeds of a 3-D solve: (x, y, z, instance) quadruples in device-readable memory.
hipError_t launch_seed3d(const SolveArgs& a, const uint32_t* seeds_xyzi, uint32_t n_seeds,
                         uint32_t tile_w, uint32_t tile_h, uint32_t tile_d, hipStream_t s);

}  // namespace ghcm
