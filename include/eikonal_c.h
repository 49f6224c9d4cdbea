/*
 * eikonal_c.h -- C entry points over the C++ API of eikonal.h.
 *
 * The reference exposes its CPU solvers and the gHCM setup/solve pair only as C++
 * functions (src/eikonal.h:40-67).  These wrappers give hosts without a C++ ABI
 * (ctypes, cgo, JNI) the same three solvers, routed through the C++ layer
 * (Field2D in, EikonalResult out) so that layer is exercised too.
 *
 * Buffers are host memory: cost[w*h] FP32 (cost = 1/speed), seeds (x,y) pairs,
 * time_out[w*h].  *duration_us receives EikonalResult::duration_total (may be NULL).
 */
#ifndef GHCM_EIKONAL_C_H
#define GHCM_EIKONAL_C_H

#include <stdint.h>

#include "ghcm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* eikonal_fmm_sequential (src/eikonal_fmm.cpp:12-71). */
int ghcm_cpu_fmm(const float* cost, uint32_t w, uint32_t h, const uint32_t* seeds_xy,
                 uint32_t n_seeds, float* time_out, double* duration_us);

/* eikonal_fim_sequential (src/eikonal_fim.cpp:13-107). */
int ghcm_cpu_fim(float minimum_change, const float* cost, uint32_t w, uint32_t h,
                 const uint32_t* seeds_xy, uint32_t n_seeds, float* time_out,
                 double* duration_us);

/* eikonal_hcm_hip_setup + eikonal_hcm_hip_ex (src/eikonal_hcm.cpp:137-529) with the
 * configuration given as HcmConfig fields: block_size, compute_unit_overload,
 * fim_minimum_change and kernel (0 systolic, 1 FIM). */
int ghcm_hcm_solve_cxx(int ordinal, uint32_t block_size, float compute_unit_overload,
                       float fim_minimum_change, int kernel, const float* cost, uint32_t w,
                       uint32_t h, const uint32_t* seeds_xy, uint32_t n_seeds,
                       float* time_out, ghcm_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
