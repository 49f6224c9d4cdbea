/*
 * oracle.h -- CPU restatement of the reference gHCM solver's arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (ghcm_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Layout follows the reference: a field is W x H, row-major, idx = x + y*W
 * (src/util_dim.h:19-22); reads outside the grid return +INF
 * (src/util_field.h:40-41).  Times are +INF for unreached cells.
 *
 * Pinning status of each function (how a plausible mistake would be caught):
 *   orc_godunov_cpu / orc_godunov_gpu / orc_godunov_safe
 *       closed-form KATs, residual invariant, monotonicity, ulp agreement
 *       between the forms (tests/test_oracle_godunov.py).
 *   orc_fmm_f32 / orc_fmm_f64
 *       1-D strips (prefix sums of cost), constant-cost closed forms,
 *       brute-force Gauss-Seidel fixed point (orc_sweep_f64, a different
 *       algorithm), obstacle/unreachable invariants (tests/test_oracle_fmm.py).
 *   orc_fim_f32
 *       agreement with the pinned FMM within the FIM threshold, plus the
 *       reference's seed-adjacency quirk reproduced on a hand-worked grid
 *       (tests/test_oracle_fim.py).
 *   orc_sweep_f64
 *       pinned against orc_fmm_f64 and the closed forms above (each pins the
 *       other: they share no code beyond the Godunov update, which is pinned
 *       separately).
 *   orc_godunov_k_f32 / orc_godunov_k_f64
 *       closed forms for K = 1, 2, 3 (a + c, the 2-D KATs, a + c/sqrt3 for three
 *       equal minima), agreement with the 2-D forms when the third axis is INF, the
 *       residual sum (T - a_i)_+^2 = c^2, monotonicity (tests/test_oracle_kd.py).
 *   orc_fmm3d_f32 / orc_fmm3d_f64
 *       1-D strips (prefix sums), D = 1 slabs equal to the 2-D FMM, constant-cost
 *       closed forms at (i,0,0), (1,1,0), (1,1,1), Euclidean <= T <= Manhattan,
 *       obstacles, and the 3-D sweeping fixed point (tests/test_oracle_kd.py).
 *   orc_sweep3d_f64
 *       pinned against orc_fmm3d_f64 and the closed forms above.
 *   orc_error_metrics
 *       hand-computed small arrays.
 */
#ifndef GHCM_ORACLE_H
#define GHCM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Godunov upwind update, reference CPU form: `eikonal_solve<2>`
 * (src/eikonal.h:121-137).  dims are sorted ascending inside.  Note the
 * reference quirk: when the larger axis value is INF the result is `cost + a`
 * WITHOUT the min against t_curr (src/eikonal.h:127-128). */
float orc_godunov_cpu(float ax, float ay, float cost, float t_curr);

/* Godunov update, reference GPU form: `eikonal_solve<2>` =
 * min(t_curr, eikonal_solve_2d(...)) (data/eikonal_common.cl:288-308). */
float orc_godunov_gpu(float ax, float ay, float cost, float t_curr);

/* Cancellation-safe form: T = m + (|d| + sqrt((sqrt2*c - |d|)(sqrt2*c + |d|)))/2,
 * m = min(a,b), d = a-b, evaluated in FP32; min'ed with t_curr. */
float orc_godunov_safe(float ax, float ay, float cost, float t_curr);

/* Same update in FP64 (no quirk; always min'ed). */
double orc_godunov_f64(double ax, double ay, double cost, double t_curr);

/* Generic K-axis update, eikonal_solve<K> (src/eikonal.h:73-118), K = 1..3, in FP32
 * and FP64: the K axis minima are sorted inside; the result is min'ed with best_prev.
 * The larger root of each quadratic is taken (the reference's release build takes the
 * smaller one, src/eikonal.h:101). */
float orc_godunov_k_f32(const float* dims, int K, float cost, float best_prev);
double orc_godunov_k_f64(const double* dims, int K, double cost, double best_prev);

/* Three-dimensional FMM: W x H x D, idx = x + W*(y + H*z), seeds as (x,y,z) triples,
 * six neighbours, the K = 3 update from accepted values (FP32 with orc_godunov_k_f32,
 * FP64 with orc_godunov_k_f64).  Returns 0, or -1 on bad arguments. */
int orc_fmm3d_f32(const float* cost, uint32_t w, uint32_t h, uint32_t d,
                  const uint32_t* seeds, uint32_t n_seeds, float* time_out);
int orc_fmm3d_f64(const double* cost, uint32_t w, uint32_t h, uint32_t d,
                  const uint32_t* seeds, uint32_t n_seeds, double* time_out);

/* Three-dimensional fast-sweeping Gauss-Seidel (eight orderings) to a fixed point, FP64.
 * Returns the rounds used, or -1 on bad arguments. */
int orc_sweep3d_f64(const double* cost, uint32_t w, uint32_t h, uint32_t d,
                    const uint32_t* seeds, uint32_t n_seeds, double* t, int max_rounds);

/* Sequential FMM (src/eikonal_fmm.cpp:12-71): lazy-deletion binary min-heap,
 * neighbours solved from accepted values only.  cost[W*H], seeds[2*n] as
 * (x,y) pairs.  time_out[W*H].  Returns 0, or -1 on bad arguments. */
int orc_fmm_f32(const float* cost, uint32_t w, uint32_t h,
                const uint32_t* seeds, uint32_t n_seeds, float* time_out);

/* Same algorithm in FP64 (the accuracy oracle). */
int orc_fmm_f64(const double* cost, uint32_t w, uint32_t h,
                const uint32_t* seeds, uint32_t n_seeds, double* time_out);

/* Sequential cell-level FIM (src/eikonal_fim.cpp:13-107): FIFO active list,
 * re-queue while |change| > minimum_change, on convergence activate any
 * neighbour that strictly decreases.  Uses the CPU Godunov form (quirk
 * included).  Returns the number of cell solves performed, or -1. */
int64_t orc_fim_f32(float minimum_change, const float* cost, uint32_t w, uint32_t h,
                    const uint32_t* seeds, uint32_t n_seeds, float* time_out);

/* Brute force: FP64 Gauss-Seidel with the 4 alternating fast-sweeping
 * orders (Zhao 2005) iterated until no cell changes.  Independent of the
 * heap-based FMM; used to pin it.  Returns the number of full rounds. */
int orc_sweep_f64(const double* cost, uint32_t w, uint32_t h,
                  const uint32_t* seeds, uint32_t n_seeds, double* time_out,
                  int max_rounds);

/* Error metrics of `got` (FP32) against `ref` (FP64):
 *   out[0] = L_inf over cells finite in both
 *   out[1] = L1 (mean absolute error) over cells finite in both
 *   out[2] = max finite ref value
 *   out[3] = number of cells whose finiteness differs
 *   out[4] = number of cells finite in both                             */
void orc_error_metrics(const float* got, const double* ref, size_t n, double out[5]);

#ifdef __cplusplus
}
#endif
#endif
