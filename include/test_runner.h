// test_runner.h -- benchmark harness of the C++ API (src/test_runner.h, src/test_runner.cpp).
//
// A job list of cost fields is solved by each solver `num_trials` times; per job the last
// result is written as <out_dir>/<solver>_<job>.blob (two size_t then W*H floats), a heat
// map .png, and a .csv of the device and total durations in microseconds:
//   <solver>,<job>, device,<d0>,<d1>,...
//   <solver>,<job>, total,<t0>,<t1>,...
// A job whose CSV exists and is not empty is skipped, as in the reference.
#pragma once

#include <functional>
#include <optional>
#include <string>
#include <vector>

#include "eikonal.h"
#include "field2d.h"

namespace ghcm {

struct Job {
    std::string    name;
    Field2D<float> field;
};

using SolverFn = std::function<EikonalResult(Field2D<float> const&, std::vector<uint2> const&)>;

constexpr float  k_fim_minimum_change = 1e-3f;   // src/test_runner.cpp:10
constexpr size_t k_num_trials         = 5;       // src/test_runner.cpp:11

void test_run(std::string const& solver_name, std::vector<Job> const& jobs,
              std::vector<uint2> const& seeds, SolverFn const& fn,
              std::string const& out_dir = "data/trials", size_t num_trials = k_num_trials);

// The reference's configuration sweep (src/test_runner.cpp:220-285): systolic with one
// worker overload and FIM with worker overloads {1, 2, 4}, each over compute-unit
// overloads {1, 2, 4, 8, 16, 32} and block sizes {8, 16, 32}, cache mode Both.
void test_ghcm(HcmDevice const& device, std::vector<Job> const& jobs,
               std::vector<uint2> const& seeds, std::string const& out_dir = "data/trials");

// ghcm_<kernel>_<cache>_cu_over_<n>_wg_over_<n>_block_sz_<n>, as src/test_runner.cpp:119-146
// (non-square tiles append _tile_<w>x<h>).
std::string hcm_config_name(HcmConfig const& config);

// Cell-wise a - b, +INF where both are +INF (src/test_runner.cpp:148-159).
Field2D<float> field_diff(Field2D<float> const& a, Field2D<float> const& b);

// Field builders of the harness (src/test_runner.cpp:26-76).
std::optional<Field2D<float>> field_speed_from_img(char const* file_path);
std::optional<Field2D<float>> field_speed_from_img_luminosity(char const* file_path);
std::optional<Field2D<float>> field_cost_from_img(char const* file_path);
Field2D<float> field_cost_from_speed(Field2D<float> f);
Field2D<float> field_cost_constant(uint2 sz);
Field2D<float> field_cost_sin_checkerboard(uint2 sz, float speed_amplitude, float checkerboard_res);
Field2D<float> field_cost_random(uint2 sz, size_t seed);
Field2D<float> field_cost_as_permeable(Field2D<float> f, float inf_replacement = 1e3f);

// The .blob written by test_run.
std::optional<Field2D<float>> load_blob(std::string const& path);

}  // namespace ghcm
