// test_runner.cpp -- benchmark harness (C++ API of test_runner.h, C-ABI ghcm_test_run).
//
// Restates the reference harness (src/test_runner.cpp): the job loop with its skip rule,
// trials, .blob / heat-map .png / .csv outputs (:163-217), the configuration sweep
// (:220-285), the configuration names (:119-146), the field diff (:148-159) and the field
// builders (:26-76).  The CSV layout and the blob layout are kept byte-compatible so the
// reference's post-processing reads them unchanged.
#include "test_runner.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>

#include "capi_common.h"
#include "field_loaders.h"
#include "ghcm_io.h"
#include "util_files.h"

namespace ghcm {

void test_run(std::string const& solver_name, std::vector<Job> const& jobs,
              std::vector<uint2> const& seeds, SolverFn const& fn, std::string const& out_dir,
              size_t num_trials) {
    const std::string base_path = out_dir + "/" + solver_name;
    if (num_trials == 0) num_trials = 1;
    for (auto&& job : jobs) {
        const std::string csv_path = base_path + "_" + job.name + ".csv";
        if (FILE* f = std::fopen(csv_path.c_str(), "rb")) {
            const auto sz = file_size(f);
            std::fclose(f);
            if (sz && *sz) {
                std::cout << "Skipping: " << solver_name << " | " << job.name << "\n";
                continue;
            }
        }
        std::vector<EikonalResult::duration_t> dur_device(num_trials), dur_total(num_trials);
        EikonalResult result;
        for (size_t i = 0; i < num_trials; ++i) {
            std::cout << "Running: " << solver_name << ", " << job.name << ", " << i << " of "
                      << num_trials << "\n";
            result = fn(job.field, seeds);
            dur_device[i] = result.duration_device;
            dur_total[i] = result.duration_total;
        }
        const uint2 sz = result.field.size();
        ghcm_blob_write((base_path + "_" + job.name + ".blob").c_str(),
                        result.field.backing_store().data(), sz.x, sz.y);
        file_write_time_field((base_path + "_" + job.name + ".png").c_str(), result.field);
        std::stringstream csv;
        csv << solver_name << "," << job.name << ", device";
        for (auto&& d : dur_device) csv << "," << d.count();
        csv << "\n";
        csv << solver_name << "," << job.name << ", total";
        for (auto&& d : dur_total) csv << "," << d.count();
        csv << "\n";
        std::ofstream(csv_path, std::ofstream::out) << csv.str();
    }
}

std::string hcm_config_name(HcmConfig const& config) {
    const char* kernel = config.kernel == HcmConfig::Kernel::FIM ? "fim"
                       : config.kernel == HcmConfig::Kernel::Sweep ? "sweep" : "systolic";
    const char* cache = "cache_both";
    switch (config.cache_mode) {
        case HcmConfig::CacheMode::None: cache = "cache_none"; break;
        case HcmConfig::CacheMode::Time: cache = "cache_time"; break;
        case HcmConfig::CacheMode::Cost: cache = "cache_cost"; break;
        case HcmConfig::CacheMode::Both: cache = "cache_both"; break;
    }
    std::stringstream name;
    name << "ghcm_" << kernel << "_" << cache << "_cu_over_" << size_t(config.compute_unit_overload)
         << "_wg_over_" << size_t(config.workers_overload) << "_block_sz_" << size_t(config.block_size);
    const uint32_t tw = config.tile_w ? config.tile_w : config.block_size;
    const uint32_t th = config.tile_h ? config.tile_h : config.block_size;
    if (tw != config.block_size || th != config.block_size) name << "_tile_" << tw << "x" << th;
    return name.str();
}

void test_ghcm(HcmDevice const& device, std::vector<Job> const& jobs, std::vector<uint2> const& seeds,
               std::string const& out_dir) {
    const uint32_t worker_overloads[] = {1, 2, 4};
    const uint32_t compute_overloads[] = {1, 2, 4, 8, 16, 32};
    const uint8_t block_sizes[] = {8, 16, 32};
    auto run_config = [&](HcmConfig const& base, bool various_workers) {
        auto subvariants = [&](HcmConfig c) {
            for (uint32_t co : compute_overloads)
                for (uint8_t bs : block_sizes) {
                    c.compute_unit_overload = (float)co;
                    c.block_size = bs;
                    c.cache_mode = HcmConfig::CacheMode::Both;
                    if (auto setup = eikonal_hcm_hip_setup(device, c)) {
                        const HcmSetup s = *setup;
                        test_run(hcm_config_name(c), jobs, seeds,
                                 [s](Field2D<float> const& f, std::vector<uint2> const& p) {
                                     return eikonal_hcm_hip_ex(s, f, p);
                                 },
                                 out_dir);
                    } else {
                        std::cerr << "\n\nError: Failed setup for: " << hcm_config_name(c) << "\n\n\n";
                    }
                }
        };
        if (various_workers) {
            for (uint32_t wo : worker_overloads) {
                HcmConfig c = base;
                c.workers_overload = (float)wo;
                subvariants(c);
            }
        } else {
            HcmConfig c = base;
            c.workers_overload = 1;
            subvariants(c);
        }
    };
    HcmConfig base{};
    base.fim_minimum_change = k_fim_minimum_change;
    HcmConfig sys = base;
    sys.kernel = HcmConfig::Kernel::Systolic;
    run_config(sys, false);
    HcmConfig fim = base;
    fim.kernel = HcmConfig::Kernel::FIM;
    run_config(fim, true);
}

Field2D<float> field_diff(Field2D<float> const& a, Field2D<float> const& b) {
    Field2D<float> diff(a.size(), INFINITY);
    const size_t n = std::min(a.size_flat(), b.size_flat());
    for (size_t i = 0; i < n; ++i) {
        const float x = a.backing_store()[i], y = b.backing_store()[i];
        const bool both_inf = std::isinf(x) && std::isinf(y);
        diff.backing_store()[i] = both_inf ? INFINITY : eikonal_time_diff(x, y);
    }
    return diff;
}

std::optional<Field2D<float>> field_speed_from_img(char const* file_path) {
    return field2d_img_monochannel(file_path);
}

std::optional<Field2D<float>> field_speed_from_img_luminosity(char const* file_path) {
    return field2d_img_rgb_luminosity(file_path);
}

Field2D<float> field_cost_from_speed(Field2D<float> f) {
    for (auto& v : f.backing_store()) v = 1 / v;
    return f;
}

std::optional<Field2D<float>> field_cost_from_img(char const* file_path) {
    auto speed = field_speed_from_img(file_path);
    if (!speed) return std::nullopt;
    return field_cost_from_speed(std::move(*speed));
}

Field2D<float> field_cost_constant(uint2 sz) { return Field2D<float>(sz, 1.f); }

namespace {
// field_cost_parameteric (src/test_runner.cpp:42-48): f(cell position / size), row-major.
template <typename F>
Field2D<float> parametric(uint2 sz, F&& f) {
    Field2D<float> field(sz);
    for (size_t i = 0; i < field.size_flat(); ++i) {
        const uint2 p = dim_pos(sz, i);
        field[i] = f((float)p.x / (float)sz.x, (float)p.y / (float)sz.y);
    }
    return field;
}
}  // namespace

Field2D<float> field_cost_sin_checkerboard(uint2 sz, float speed_amplitude, float checkerboard_res) {
    const float k = float(checkerboard_res * 2 * M_PI);
    return parametric(sz, [&](float x, float y) {
        const float vx = std::sin(k * x), vy = std::sin(k * y);
        return float(1 + speed_amplitude * vx * vy);
    });
}

Field2D<float> field_cost_random(uint2 sz, size_t seed) {
    // std::default_random_engine of the reference's standard library: minstd_rand0.
    using rnd_t = std::minstd_rand0;
    rnd_t gen((rnd_t::result_type)seed);
    return parametric(sz, [&](float, float) {
        const float n = float(gen() - rnd_t::min()) / rnd_t::max();
        return 1 + n * 5;
    });
}

Field2D<float> field_cost_as_permeable(Field2D<float> f, float inf_replacement) {
    for (auto&& x : f.backing_store()) x = std::min(x, inf_replacement);
    return f;
}

std::optional<Field2D<float>> load_blob(std::string const& path) {
    ghcm_image img{};
    if (ghcm_blob_read(path.c_str(), &img) != GHCM_OK) return std::nullopt;
    Field2D<float> f(uint2(img.w, img.h), img.data);
    ghcm_image_free(&img);
    return f;
}

}  // namespace ghcm

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------

extern "C" int ghcm_test_run(const char* solver, ghcm_ctx* ctx, const char* solver_name,
                             const char* job_name, const float* cost, uint32_t w, uint32_t h,
                             const uint32_t* seeds_xy, uint32_t n_seeds, const char* out_dir,
                             uint32_t trials, float fim_minimum_change) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!solver || !solver_name || !job_name || !cost || !out_dir || !w || !h || (n_seeds && !seeds_xy))
            return ghcm_capi::fail(GHCM_ERR_INVALID_ARG, "null or empty argument");
        const std::string kind = solver;
        ghcm::SolverFn fn;
        int status = GHCM_OK;
        if (kind == "fmm") {
            fn = ghcm::eikonal_fmm_sequential;
        } else if (kind == "fim") {
            fn = [fim_minimum_change](ghcm::Field2D<float> const& f, std::vector<ghcm::uint2> const& s) {
                return ghcm::eikonal_fim_sequential(fim_minimum_change, f, s);
            };
        } else if (kind == "hcm") {
            if (!ctx) return ghcm_capi::fail(GHCM_ERR_INVALID_ARG, "the hcm solver needs a context");
            fn = [ctx, &status](ghcm::Field2D<float> const& f, std::vector<ghcm::uint2> const& s) {
                ghcm::EikonalResult r;
                std::vector<uint32_t> xy;
                for (auto&& p : s) { xy.push_back(p.x); xy.push_back(p.y); }
                r.field = ghcm::Field2D<float>(f.size());
                const int rc = ghcm_solve(ctx, f.backing_store().data(), GHCM_FIELD_COST_F32, f.size().x,
                                          f.size().y, xy.data(), (uint32_t)s.size(),
                                          r.field.backing_store().data(), &r.stats);
                if (rc != GHCM_OK) status = rc;
                r.status = rc;
                r.duration_total = ghcm::EikonalResult::duration_t((long long)r.stats.duration_total_us);
                r.duration_device = ghcm::EikonalResult::duration_t((long long)r.stats.duration_device_us);
                return r;
            };
        } else {
            return ghcm_capi::fail(GHCM_ERR_INVALID_ARG, "unknown solver %s (fmm, fim or hcm)", solver);
        }
        ghcm::Field2D<float> field(ghcm::uint2(w, h), cost);
        std::vector<ghcm::uint2> seeds;
        for (uint32_t i = 0; i < n_seeds; ++i) seeds.emplace_back(seeds_xy[2 * i], seeds_xy[2 * i + 1]);
        ghcm::test_run(solver_name, {ghcm::Job{job_name, std::move(field)}}, seeds, fn, out_dir, trials);
        return status;
    });
}

extern "C" int ghcm_field_cost_sin_checkerboard(uint32_t w, uint32_t h, float amp, float res, float* out) {
    if (!out || !w || !h) return ghcm_capi::fail(GHCM_ERR_INVALID_ARG, "bad arguments");
    const auto f = ghcm::field_cost_sin_checkerboard(ghcm::uint2(w, h), amp, res);
    std::copy(f.backing_store().begin(), f.backing_store().end(), out);
    return GHCM_OK;
}

extern "C" int ghcm_field_cost_random(uint32_t w, uint32_t h, uint64_t seed, float* out) {
    if (!out || !w || !h) return ghcm_capi::fail(GHCM_ERR_INVALID_ARG, "bad arguments");
    const auto f = ghcm::field_cost_random(ghcm::uint2(w, h), (size_t)seed);
    std::copy(f.backing_store().begin(), f.backing_store().end(), out);
    return GHCM_OK;
}
