// autotune.cpp -- timing-based choice of tile shape, n, in-tile rounds and launch mode.
//
// The reference explores its configuration grid by running every (compute-unit overload,
// block size, workers overload) combination over its jobs and writing CSV timings
// (src/test_runner.cpp:220-285); choosing among them is left to the reader.  Here the
// same idea runs inside the library: every candidate solves the caller's workload `reps`
// times, the best device time counts, and the fastest converged candidate becomes the
// context configuration.
#include <cmath>
#include <cstring>
#include <initializer_list>

#include "capi_common.h"
#include "ghcm.h"

using ghcm_capi::fail;

namespace {

void put(ghcm_tune_candidate* out, uint32_t cap, uint32_t* n, uint32_t tw, uint32_t th, int kernel,
         uint32_t n_active, uint32_t inner, uint32_t mode) {
    if (*n < cap) {
        ghcm_tune_candidate& c = out[*n];
        std::memset(&c, 0, sizeof c);
        c.tile_w = tw;
        c.tile_h = th;
        c.kernel = kernel;
        c.n_active = n_active;
        c.inner_iter_cap = inner;
        c.launch_mode = mode;
        c.device_us = INFINITY;
    }
    ++*n;
}

}  // namespace

extern "C" int ghcm_tune_default_space(int cus, ghcm_tune_candidate* out, uint32_t cap,
                                       uint32_t* n_out) {
    if (!n_out || (cap && !out) || cus <= 0)
        return fail(GHCM_ERR_INVALID_ARG, "tune space: null output or compute_units <= 0");
    uint32_t n = 0;
    const uint32_t mult[5] = {1, 2, 4, 16, 64};
    const uint32_t small[3][2] = {{8, 8}, {16, 16}, {32, 8}};
    const uint32_t large[3][2] = {{64, 64}, {64, 32}, {32, 32}};
    for (uint32_t mode : {0u, 2u})
        for (uint32_t m : mult) {
            const uint32_t na = m * (uint32_t)cus;
            for (const auto& t : large)
                for (uint32_t inner : {1u, 2u, 0u})
                    put(out, cap, &n, t[0], t[1], GHCM_KERNEL_SWEEP, na, inner, mode);
            for (const auto& t : small) {
                for (uint32_t inner : {0u, 8u})
                    put(out, cap, &n, t[0], t[1], GHCM_KERNEL_SYSTOLIC, na, inner, mode);
                put(out, cap, &n, t[0], t[1], GHCM_KERNEL_FIM, na, 0u, mode);
            }
        }
    // The sweep kernel's other wave loops: one persistent cooperative launch (1),
    // boundary-free waves (4, every queued tile each wave, n sets the grid) and boundary-free
    // waves with the top-n cut made by the tile launch (5; 64x64 and 64x32 tiles only, the
    // others would run as 0).
    for (uint32_t mode : {1u, 4u, 5u})
        for (uint32_t m : mult) {
            const uint32_t na = m * (uint32_t)cus;
            for (const auto& t : large) {
                if (mode == 5u && t[0] != 64u) continue;
                for (uint32_t inner : {1u, 0u})
                    put(out, cap, &n, t[0], t[1], GHCM_KERNEL_SWEEP, na, inner, mode);
            }
        }
    *n_out = n < cap ? n : cap;
    if (n > cap) return fail(GHCM_ERR_INVALID_ARG, "tune space has %u candidates, capacity %u", n, cap);
    return GHCM_OK;
}

extern "C" int ghcm_autotune(ghcm_ctx* ctx, int kind, const ghcm_instance* inst, uint32_t n_inst,
                             ghcm_tune_candidate* cands, uint32_t n_cands, uint32_t reps,
                             uint32_t* best) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!ctx || !inst || n_inst == 0 || !cands || n_cands == 0)
            return fail(GHCM_ERR_INVALID_ARG, "autotune: null argument or empty instance/candidate list");
        ghcm_config base;
        int rc = ghcm_ctx_get_config(ctx, &base);
        if (rc) return rc;
        if (reps == 0) reps = 1;
        uint32_t best_i = n_cands;
        for (uint32_t i = 0; i < n_cands; ++i) {
            ghcm_tune_candidate& c = cands[i];
            c.device_us = INFINITY;
            c.converged = 0;
            c.tile_visits = 0;
            ghcm_config cfg = base;
            cfg.tile_w = c.tile_w;
            cfg.tile_h = c.tile_h;
            cfg.kernel = c.kernel;
            cfg.n_active = c.n_active;
            cfg.inner_iter_cap = c.inner_iter_cap;
            cfg.launch_mode = c.launch_mode;
            cfg.profile_kernels = 0;
            c.status = ghcm_ctx_set_config(ctx, &cfg);
            if (c.status != GHCM_OK) continue;
            for (uint32_t r = 0; r < reps; ++r) {
                ghcm_stats st;
                c.status = ghcm_solve_batch(ctx, kind, inst, n_inst, &st);
                if (c.status != GHCM_OK) break;
                c.converged = st.converged;
                c.tile_visits = st.tile_visits;
                if (st.duration_device_us < c.device_us) c.device_us = st.duration_device_us;
            }
            if (c.status == GHCM_OK && c.converged &&
                (best_i == n_cands || c.device_us < cands[best_i].device_us))
                best_i = i;
        }
        if (best_i == n_cands) {
            (void)ghcm_ctx_set_config(ctx, &base);
            return fail(GHCM_ERR_NOT_CONVERGED, "autotune: no candidate converged");
        }
        ghcm_config cfg = base;
        cfg.tile_w = cands[best_i].tile_w;
        cfg.tile_h = cands[best_i].tile_h;
        cfg.kernel = cands[best_i].kernel;
        cfg.n_active = cands[best_i].n_active;
        cfg.inner_iter_cap = cands[best_i].inner_iter_cap;
        cfg.launch_mode = cands[best_i].launch_mode;
        rc = ghcm_ctx_set_config(ctx, &cfg);
        if (best) *best = best_i;
        return rc;
    });
}
