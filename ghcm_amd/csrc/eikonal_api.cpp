// eikonal_api.cpp -- C++ layer of eikonal.h over the C-ABI, plus its C wrappers.
//
// eikonal_hcm_hip_setup / _ex / eikonal_hcm_hip mirror eikonal_hcm_opencl_setup /
// _ex / eikonal_hcm_opencl (src/eikonal_hcm.cpp:137-184, 186-529): setup validates the
// configuration and creates the device context; _ex solves one cost field and returns
// the time field by value with total/device durations.
#include <iostream>

#include "capi_common.h"
#include "eikonal.h"
#include "eikonal_c.h"

namespace ghcm {

ghcm_config HcmConfig::to_c() const {
    ghcm_config c;
    ghcm_config_default(&c);
    const uint32_t b = block_size ? block_size : 1u;
    c.tile_w = tile_w ? tile_w : b;
    c.tile_h = tile_h ? tile_h : b;
    c.workers_overload = workers_overload;
    c.compute_unit_overload = compute_unit_overload;
    c.n_active = n_active;
    c.min_change = fim_minimum_change;
    c.cache_mode = (int)cache_mode;
    c.kernel = kernel == Kernel::FIM ? GHCM_KERNEL_FIM
             : kernel == Kernel::Sweep ? GHCM_KERNEL_SWEEP : GHCM_KERNEL_SYSTOLIC;
    c.scheduler = scheduler == Scheduler::HostPQ ? GHCM_SCHED_HOST_PQ : GHCM_SCHED_DEVICE_TOPN;
    c.inner_iter_cap = inner_iter_cap;
    c.launch_mode = launch_mode;
    return c;
}

// Same fields and wording as src/eikonal_hcm.cpp:116-135, plus the added knobs.
std::ostream& operator<<(std::ostream& os, HcmConfig const& c) {
    const char* mode = "Time & Cost";
    switch (c.cache_mode) {
        case HcmConfig::CacheMode::None: mode = "None"; break;
        case HcmConfig::CacheMode::Time: mode = "Time"; break;
        case HcmConfig::CacheMode::Cost: mode = "Cost"; break;
        case HcmConfig::CacheMode::Both: mode = "Time & Cost"; break;
    }
    const ghcm_config cc = c.to_c();
    os << "FIM Threshold         : " << c.fim_minimum_change
       << "\nBlock Size            : " << (size_t)c.block_size
       << "\nWorker Overload       : " << c.workers_overload
       << "\nCompute Unit Overload : " << c.compute_unit_overload
       << "\nCache Mode            : " << mode
       << "\nTile                  : " << cc.tile_w << "x" << cc.tile_h
       << "\nKernel                : " << (c.kernel == HcmConfig::Kernel::FIM ? "FIM"
                                           : c.kernel == HcmConfig::Kernel::Sweep ? "Sweep" : "Systolic")
       << "\nScheduler             : " << (c.scheduler == HcmConfig::Scheduler::HostPQ ? "host PQ" : "device top-n")
       << "\n";
    return os;
}

std::optional<HcmSetup> eikonal_hcm_hip_setup(HcmDevice const& device, HcmConfig const& config) {
    const ghcm_config c = config.to_c();
    ghcm_ctx* raw = nullptr;
    if (ghcm_ctx_create(device.ordinal, &c, &raw) != GHCM_OK) {
        std::cout << "gHCM setup failed: " << ghcm_last_error() << "\n";
        return std::nullopt;
    }
    HcmSetup s;
    s.config = config;
    s.device = device;
    s.ctx = std::shared_ptr<ghcm_ctx>(raw, [](ghcm_ctx* p) { ghcm_ctx_destroy(p); });
    return s;
}

EikonalResult eikonal_hcm_hip_ex(HcmSetup info, Field2D<float> const& field_cost,
                                 std::vector<uint2> const& seed_points) {
    EikonalResult r;
    const uint2 sz = field_cost.size();
    r.field = Field2D<float>(sz, INFINITY);
    if (!info.ctx) {
        r.status = GHCM_ERR_INVALID_ARG;
        return r;
    }
    std::vector<uint32_t> seeds;
    seeds.reserve(seed_points.size() * 2);
    for (const uint2& p : seed_points) {
        seeds.push_back(p.x);
        seeds.push_back(p.y);
    }
    r.status = ghcm_solve(info.ctx.get(), field_cost.backing_store().data(), GHCM_FIELD_COST_F32,
                          sz.x, sz.y, seeds.data(), (uint32_t)seed_points.size(),
                          r.field.backing_store().data(), &r.stats);
    r.duration_total = EikonalResult::duration_t((long long)r.stats.duration_total_us);
    r.duration_device = EikonalResult::duration_t((long long)r.stats.duration_device_us);
    return r;
}

EikonalResult eikonal_hcm_hip(HcmDevice const& device, HcmConfig const& config,
                              Field2D<float> const& field_cost,
                              std::vector<uint2> const& seed_points) {
    auto setup = eikonal_hcm_hip_setup(device, config);
    if (!setup) {
        EikonalResult r;
        r.field = Field2D<float>(field_cost.size(), INFINITY);
        r.status = GHCM_ERR_INVALID_ARG;
        return r;
    }
    return eikonal_hcm_hip_ex(*setup, field_cost, seed_points);
}

}  // namespace ghcm

// ---- C wrappers ----------------------------------------------------------------------

namespace {

bool to_inputs(const float* cost, uint32_t w, uint32_t h, const uint32_t* seeds, uint32_t n,
               ghcm::Field2D<float>* f, std::vector<ghcm::uint2>* s) {
    if (!cost || w == 0 || h == 0 || (n && !seeds)) return false;
    *f = ghcm::Field2D<float>(ghcm::uint2(w, h), cost);
    s->clear();
    for (uint32_t i = 0; i < n; ++i) {
        if (seeds[2 * i] >= w || seeds[2 * i + 1] >= h) return false;
        s->emplace_back(seeds[2 * i], seeds[2 * i + 1]);
    }
    return true;
}

int copy_out(const ghcm::EikonalResult& r, float* out, double* us) {
    if (r.status != GHCM_OK) return r.status;
    std::copy(r.field.backing_store().begin(), r.field.backing_store().end(), out);
    if (us) *us = (double)r.duration_total.count();
    return GHCM_OK;
}

}  // namespace

extern "C" int ghcm_cpu_fmm(const float* cost, uint32_t w, uint32_t h, const uint32_t* seeds,
                            uint32_t n, float* out, double* us) {
    return ghcm_capi::no_throw([&]() -> int {
        ghcm::Field2D<float> f;
        std::vector<ghcm::uint2> s;
        if (!out || !to_inputs(cost, w, h, seeds, n, &f, &s)) return GHCM_ERR_INVALID_ARG;
        return copy_out(ghcm::eikonal_fmm_sequential(f, s), out, us);
    });
}

extern "C" int ghcm_cpu_fim(float eps, const float* cost, uint32_t w, uint32_t h,
                            const uint32_t* seeds, uint32_t n, float* out, double* us) {
    return ghcm_capi::no_throw([&]() -> int {
        ghcm::Field2D<float> f;
        std::vector<ghcm::uint2> s;
        if (!out || !to_inputs(cost, w, h, seeds, n, &f, &s)) return GHCM_ERR_INVALID_ARG;
        return copy_out(ghcm::eikonal_fim_sequential(eps, f, s), out, us);
    });
}

extern "C" int ghcm_hcm_solve_cxx(int ordinal, uint32_t block_size, float cu_overload, float eps,
                                  int kernel, const float* cost, uint32_t w, uint32_t h,
                                  const uint32_t* seeds, uint32_t n, float* out, ghcm_stats* stats) {
    return ghcm_capi::no_throw([&]() -> int {
        ghcm::Field2D<float> f;
        std::vector<ghcm::uint2> s;
        if (!out || block_size == 0 || block_size > 255 || !to_inputs(cost, w, h, seeds, n, &f, &s))
            return GHCM_ERR_INVALID_ARG;
        ghcm::HcmConfig c;
        c.block_size = (uint8_t)block_size;
        c.compute_unit_overload = cu_overload;
        c.fim_minimum_change = eps;
        c.kernel = kernel == 1 ? ghcm::HcmConfig::Kernel::FIM
                 : kernel == 2 ? ghcm::HcmConfig::Kernel::Sweep : ghcm::HcmConfig::Kernel::Systolic;
        auto setup = ghcm::eikonal_hcm_hip_setup(ghcm::HcmDevice{ordinal}, c);
        if (!setup) return GHCM_ERR_INVALID_ARG;
        ghcm::EikonalResult r = ghcm::eikonal_hcm_hip_ex(*setup, f, s);
        if (stats) *stats = r.stats;
        if (r.status != GHCM_OK) return r.status;
        std::copy(r.field.backing_store().begin(), r.field.backing_store().end(), out);
        return GHCM_OK;
    });
}
