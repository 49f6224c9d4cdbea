// A caller written the way INTEGRATION.md section 1 tells a maintainer of the reference's
// harness to switch: src/eikonal.h replaced by include/eikonal.h, the solver bound into a
// SolverFn with std::bind (src/test_runner.cpp:117, 243-246), and libghcm.so linked.  It is
// compiled by g++ outside the library and run by tests/test_dropin_cxx.py.
//
// usage: dropin_caller <solver: fmm|fim|hcm|hcm_sweep> <cost.bin> <w> <h> <out.bin> x0 y0 [x1 y1 ...]
//   cost.bin: w*h float32 costs, row-major; out.bin receives w*h float32 times.
//   exit status 0 on success, 2 on bad usage, 3 if setup failed, 4 if the solve failed.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "eikonal.h"
#include "test_runner.h"

using namespace ghcm;

int main(int argc, char** argv) {
    if (argc < 8 || (argc - 6) % 2 != 0) {
        std::fprintf(stderr, "usage: %s solver cost.bin w h out.bin x0 y0 [x y ...]\n", argv[0]);
        return 2;
    }
    const std::string solver = argv[1];
    const uint2 sz{(uint32_t)std::atoi(argv[3]), (uint32_t)std::atoi(argv[4])};
    Field2D<float> field_cost(sz);
    {
        FILE* f = std::fopen(argv[2], "rb");
        if (!f || std::fread(field_cost.backing_store().data(), 4, field_cost.size_flat(), f) !=
                      field_cost.size_flat()) {
            std::fprintf(stderr, "cannot read %s\n", argv[2]);
            return 2;
        }
        std::fclose(f);
    }
    std::vector<uint2> seeds;
    for (int i = 6; i + 1 < argc; i += 2) seeds.push_back({(uint32_t)std::atoi(argv[i]), (uint32_t)std::atoi(argv[i + 1])});

    SolverFn fn;
    if (solver == "fmm") {
        fn = eikonal_fmm_sequential;
    } else if (solver == "fim") {
        fn = std::bind(eikonal_fim_sequential, k_fim_minimum_change, std::placeholders::_1,
                       std::placeholders::_2);
    } else if (solver == "hcm" || solver == "hcm_sweep") {
        HcmConfig c;                   // the reference's defaults (src/eikonal.h:12-27)
        if (solver == "hcm") {
            c.block_size = 16;
            c.compute_unit_overload = 4;
            c.kernel = HcmConfig::Kernel::Systolic;
        } else {
            c.tile_w = c.tile_h = 64;
            c.kernel = HcmConfig::Kernel::Sweep;
            c.n_active = 512;
            c.inner_iter_cap = 1;
        }
        auto setup = eikonal_hcm_opencl_setup(HcmDevice{0}, c);
        if (!setup) {
            std::fprintf(stderr, "setup failed: %s\n", ghcm_last_error());
            return 3;
        }
        fn = std::bind(eikonal_hcm_opencl_ex, *setup, std::placeholders::_1, std::placeholders::_2);
    } else {
        std::fprintf(stderr, "unknown solver %s\n", solver.c_str());
        return 2;
    }
    EikonalResult r = fn(field_cost, seeds);
    if (r.status != GHCM_OK) {
        std::fprintf(stderr, "solve failed: %s\n", ghcm_last_error());
        return 4;
    }
    FILE* f = std::fopen(argv[5], "wb");
    if (!f || std::fwrite(r.field.backing_store().data(), 4, r.field.size_flat(), f) != r.field.size_flat())
        return 2;
    std::fclose(f);
    std::printf("%s: %ux%u, total %lld us, device %lld us\n", solver.c_str(), sz.x, sz.y,
                (long long)r.duration_total.count(), (long long)r.duration_device.count());
    return 0;
}
