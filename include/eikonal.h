// eikonal.h -- C++ solver API, the drop-in counterpart of the reference's src/eikonal.h.
//
// Same entry points, argument meaning and result type as src/eikonal.h:12-143:
//   HcmConfig                 src/eikonal.h:12-27  (plus MI355X knobs, defaults unchanged)
//   EikonalResult             src/eikonal.h:31-36  (plus solver counters)
//   eikonal_fmm_sequential    src/eikonal.h:40-41
//   eikonal_fim_sequential    src/eikonal.h:44-46
//   eikonal_hcm_opencl        src/eikonal.h:48-51  -> eikonal_hcm_hip (alias kept)
//   HcmSetup                  src/eikonal.h:53-60  (HIP context instead of CL handles)
//   eikonal_hcm_opencl_setup  src/eikonal.h:62-63  -> eikonal_hcm_hip_setup (alias kept)
//   eikonal_hcm_opencl_ex     src/eikonal.h:65-67  -> eikonal_hcm_hip_ex (alias kept)
//   eikonal_solve<K>          src/eikonal.h:73-137 (CPU Godunov update)
//   eikonal_time_diff         src/eikonal.h:142-143
// The input field is a per-cell COST (1/speed), as in the reference.  The device is
// a HIP ordinal (HcmDevice) instead of a cl::Device.  Failures return an empty
// optional / a result whose `status` is non-zero; nothing calls exit().
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <memory>
#include <optional>
#include <ostream>
#include <vector>

#include "field2d.h"
#include "ghcm.h"

namespace ghcm {

struct HcmConfig {
    enum class Kernel { Systolic, FIM, Sweep };   // Sweep: MI355X fast-sweeping tile solver
    enum class CacheMode { None, Time, Cost, Both };
    enum class Scheduler { DeviceTopN, HostPQ };

    // `block_size` by `block_size` tiles (8, 16 or 32).
    uint8_t   block_size            = 8;
    // Reference knob: work-group size multiple of 64.  A wave64 solves one tile here.
    float     workers_overload      = 1.f;
    // n = compute_unit_overload x #CUs tiles per wave.
    float     compute_unit_overload = 1.f;
    // A cell is converged when it changes by less than this threshold.
    float     fim_minimum_change    = 1e-5f;
    CacheMode cache_mode            = CacheMode::Both;
    Kernel    kernel                = Kernel::Systolic;
    // MI355X extensions (zero = derived from the fields above).
    uint32_t  tile_w = 0, tile_h = 0;   // non-square tiles, e.g. 32x8
    uint32_t  n_active = 0;             // overrides compute_unit_overload
    uint32_t  inner_iter_cap = 0;       // in-tile passes per visit (0 = until converged)
    Scheduler scheduler = Scheduler::DeviceTopN;
    uint32_t  launch_mode = 0;          // 0 per-wave launches, 1 persistent, 2 HIP graph

    ghcm_config to_c() const;
};
std::ostream& operator<<(std::ostream& os, HcmConfig const& c);

struct EikonalResult {
    using duration_t = std::chrono::microseconds;

    Field2D<float> field;
    duration_t     duration_total{0}, duration_device{0};
    // additions: solver status and counters (zero for the CPU solvers)
    int            status = GHCM_OK;
    ghcm_stats     stats{};
};

// Sequential Fast Marching on the CPU (src/eikonal_fmm.cpp:12-71).
EikonalResult eikonal_fmm_sequential(Field2D<float> const& field_cost,
                                     std::vector<uint2> const& seed_points);

// Sequential cell-level Fast Iterative Method on the CPU (src/eikonal_fim.cpp:13-107).
EikonalResult eikonal_fim_sequential(float minimum_change, Field2D<float> const& field_cost,
                                     std::vector<uint2> const& seed_points);

// A HIP device ordinal; replaces cl::Device.
struct HcmDevice {
    int ordinal = 0;
};

// Replaces {device, ctx, cmd_queue, program, program_dbg}: kernels are compiled
// ahead of time, so setup owns one context (HIP stream + device scratch).
struct HcmSetup {
    HcmConfig                 config;
    HcmDevice                 device;
    std::shared_ptr<ghcm_ctx> ctx;
};

std::optional<HcmSetup> eikonal_hcm_hip_setup(HcmDevice const& device, HcmConfig const& config);

EikonalResult eikonal_hcm_hip_ex(HcmSetup info, Field2D<float> const& field_cost,
                                 std::vector<uint2> const& seed_points);

EikonalResult eikonal_hcm_hip(HcmDevice const& device, HcmConfig const& config,
                              Field2D<float> const& field_cost,
                              std::vector<uint2> const& seed_points);

// Source-compatible spellings of the reference's names.
inline std::optional<HcmSetup> eikonal_hcm_opencl_setup(HcmDevice const& d, HcmConfig const& c) {
    return eikonal_hcm_hip_setup(d, c);
}
inline EikonalResult eikonal_hcm_opencl_ex(HcmSetup info, Field2D<float> const& f,
                                           std::vector<uint2> const& s) {
    return eikonal_hcm_hip_ex(std::move(info), f, s);
}
inline EikonalResult eikonal_hcm_opencl(HcmDevice const& d, HcmConfig const& c,
                                        Field2D<float> const& f, std::vector<uint2> const& s) {
    return eikonal_hcm_hip(d, c, f, s);
}

// ---- CPU Godunov update (src/eikonal.h:70-137) ----
typedef float EikonalAxisMin;

// Generic K-axis update: sorted axis minima; solve sum_j (T - a_j)^2 = cost^2 over the
// first i+1 axes while the root exceeds the next axis value.
template <size_t K>
float eikonal_solve(EikonalAxisMin const (&dim_sorted)[K], float const cost, float const best_prev) {
    static_assert(K >= 1 && K <= 3, "1 to 3 axes");
    float b = 0, c = -(cost * cost), time_best = best_prev;
    for (size_t i = 0; i < K; ++i) {
        const float ta = dim_sorted[i];
        if (time_best <= ta) break;
        if (std::isinf(ta)) return time_best;
        b += -2 * ta;
        c += ta * ta;
        if (i == 0) {
            time_best = std::min(time_best, ta + cost);
        } else {
            const float a = float(i) + 1;
            const float quad = b * b - 4 * a * c;
            const float t = (-b + std::sqrt(std::max(quad, 0.f))) / (2 * a);
            time_best = std::min(t, time_best);
        }
    }
    return time_best;
}

// Two-axis update in the reference's CPU form, including its behaviour when the
// larger axis value is INF: `cost + a` is returned without the min against the
// current time (src/eikonal.h:127-128).
template <>
inline float eikonal_solve<2>(EikonalAxisMin const (&dim_sorted)[2], float const cost,
                              float const time_curr) {
    const float dX = dim_sorted[0], dY = dim_sorted[1];
    if (std::isinf(dX)) return cost + dY;
    if (std::isinf(dY)) return cost + dX;
    const float dXY = dX - dY;
    if (cost < std::fabs(dXY)) return std::min(time_curr, std::min(dX, dY) + cost);
    const float q = cost * cost * 2 - dXY * dXY;
    const float p = (dX + dY + std::sqrt(q)) / 2;
    return std::min(time_curr, p);
}

// a - b, with inf - inf = 0 instead of NaN (src/eikonal.h:142-143).
inline float eikonal_time_diff(float a, float b) { return (a == b) ? 0 : (a - b); }

}  // namespace ghcm
