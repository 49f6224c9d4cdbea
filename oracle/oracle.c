/*
 * oracle.c -- CPU restatement of the reference gHCM arithmetic (test infrastructure).
 *
 * This file is the checker for the HIP path.  It is compiled by oracle/Makefile into
 * oracle/liboracle.so and used only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Build flags: -O2 -ffp-contract=off (no FMA
 * contraction, like the reference's default x86-64 build).
 *
 * Reference files followed (paths relative to the reference checkout):
 *   src/eikonal.h:121-137         CPU Godunov update eikonal_solve<2>
 *   data/eikonal_common.cl:288-308 GPU Godunov update (always min'ed)
 *   src/eikonal_fmm.cpp:12-71     sequential FMM
 *   src/eikonal_fim.cpp:13-107    sequential FIM
 *   src/util_field.h:40-41        out-of-grid reads are +INF
 *   src/util_dim.h:19-22          idx = x + y*W
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Godunov updates                                                            */
/* ------------------------------------------------------------------------- */

/* src/eikonal.h:121-137.  Axis values are sorted so dX <= dY first
 * (std::sort in src/eikonal_fmm.cpp:32 / src/eikonal_fim.cpp:41). */
float orc_godunov_cpu(float ax, float ay, float cost, float t_curr) {
    float dX = ax < ay ? ax : ay;
    float dY = ax < ay ? ay : ax;
    if (isinf(dX)) return cost + dY;
    if (isinf(dY)) return cost + dX;           /* quirk: no min with t_curr */
    float dXY = dX - dY;
    if (cost < fabsf(dXY)) {
        float m = (dX < dY ? dX : dY) + cost;
        return t_curr < m ? t_curr : m;
    }
    float q = cost * cost * 2 - dXY * dXY;
    float p = (dX + dY + sqrtf(q)) / 2;
    return t_curr < p ? t_curr : p;
}

/* data/eikonal_common.cl:288-308: eikonal_solve_2d then min with t_curr. */
float orc_godunov_gpu(float ax, float ay, float cost, float t_curr) {
    float dX = ax < ay ? ax : ay;
    float dY = ax < ay ? ay : ax;
    float r;
    if (isinf(dX))      r = cost + dY;
    else if (isinf(dY)) r = cost + dX;
    else {
        float dXY = dX - dY;
        if (cost < fabsf(dXY)) r = (dX < dY ? dX : dY) + cost;
        else {
            float q = cost * cost * 2 - dXY * dXY;
            r = (dX + dY + sqrtf(q)) / 2;
        }
    }
    return t_curr < r ? t_curr : r;
}

/* Factored discriminant: 2c^2 - d^2 = (sqrt2 c - |d|)(sqrt2 c + |d|), and the
 * result formed as min(a,b) + increment so the large common part is added once. */
float orc_godunov_safe(float ax, float ay, float cost, float t_curr) {
    float m = ax < ay ? ax : ay;
    float M = ax < ay ? ay : ax;
    float r;
    if (isinf(m))      r = cost + M;
    else if (isinf(M)) r = cost + m;
    else {
        float d = M - m;                    /* >= 0 */
        if (cost < d) r = m + cost;
        else {
            float s = 1.41421356237309515f * cost;
            float q = (s - d) * (s + d);
            r = m + 0.5f * (d + sqrtf(q));
        }
    }
    return t_curr < r ? t_curr : r;
}

double orc_godunov_f64(double ax, double ay, double cost, double t_curr) {
    double m = ax < ay ? ax : ay;
    double M = ax < ay ? ay : ax;
    double r;
    if (isinf(m))      r = cost + M;
    else if (isinf(M)) r = cost + m;
    else {
        double d = M - m;
        if (cost < d) r = m + cost;
        else r = (m + M + sqrt(2.0 * cost * cost - d * d)) / 2.0;
    }
    return t_curr < r ? t_curr : r;
}

/* ------------------------------------------------------------------------- */
/* Binary min-heap of (key, cell index), restating std::priority_queue with   */
/* the inverted comparator of src/eikonal_fmm.cpp:14-19.                       */
/* ------------------------------------------------------------------------- */

#define DEFINE_HEAP(NAME, KEY_T)                                                   \
typedef struct { KEY_T key; uint32_t idx; } NAME##_ent;                             \
typedef struct { NAME##_ent* a; size_t n, cap; } NAME;                               \
static int NAME##_push(NAME* h, KEY_T key, uint32_t idx) {                          \
    if (h->n == h->cap) {                                                           \
        size_t nc = h->cap ? h->cap * 2 : 1024;                                     \
        NAME##_ent* na = (NAME##_ent*)realloc(h->a, nc * sizeof(NAME##_ent));       \
        if (!na) return -1;                                                         \
        h->a = na; h->cap = nc;                                                     \
    }                                                                               \
    size_t i = h->n++;                                                              \
    while (i > 0) {                                                                 \
        size_t p = (i - 1) / 2;                                                     \
        if (!(key < h->a[p].key)) break;                                            \
        h->a[i] = h->a[p]; i = p;                                                   \
    }                                                                               \
    h->a[i].key = key; h->a[i].idx = idx;                                           \
    return 0;                                                                       \
}                                                                                   \
static NAME##_ent NAME##_pop(NAME* h) {                                             \
    NAME##_ent top = h->a[0];                                                       \
    NAME##_ent last = h->a[--h->n];                                                 \
    size_t i = 0, n = h->n;                                                         \
    for (;;) {                                                                      \
        size_t l = 2 * i + 1;                                                       \
        if (l >= n) break;                                                          \
        size_t c = (l + 1 < n && h->a[l + 1].key < h->a[l].key) ? l + 1 : l;        \
        if (!(h->a[c].key < last.key)) break;                                       \
        h->a[i] = h->a[c]; i = c;                                                   \
    }                                                                               \
    if (n) h->a[i] = last;                                                          \
    return top;                                                                     \
}

DEFINE_HEAP(heap32, float)
DEFINE_HEAP(heap64, double)

/* ------------------------------------------------------------------------- */
/* FMM                                                                          */
/* ------------------------------------------------------------------------- */

/* Neighbour visit order of src/eikonal_fmm.cpp:61-62 (add_neighbour_axis(x) then
 * (y), each as -axis then +axis). */
static const int k_off_x[4] = { -1, 1, 0, 0 };
static const int k_off_y[4] = { 0, 0, -1, 1 };

#define FMM_BODY(KEY_T, HEAP, GODUNOV, INF_T)                                            \
    if (!cost || !time_out || w == 0 || h == 0 || (n_seeds && !seeds)) return -1;        \
    size_t N = (size_t)w * h;                                                            \
    for (size_t i = 0; i < N; ++i) time_out[i] = INF_T;                                  \
    HEAP hp = { 0, 0, 0 };                                                               \
    for (uint32_t s = 0; s < n_seeds; ++s) {                                             \
        uint32_t x = seeds[2 * s], y = seeds[2 * s + 1];                                  \
        if (x >= w || y >= h) { free(hp.a); return -1; }                                 \
        if (HEAP##_push(&hp, (KEY_T)0, x + y * w)) { free(hp.a); return -1; }            \
    }                                                                                    \
    while (hp.n) {                                                                       \
        HEAP##_ent cur = HEAP##_pop(&hp);                                                \
        if (time_out[cur.idx] <= cur.key) continue;                                      \
        time_out[cur.idx] = cur.key;                                                     \
        int cx = (int)(cur.idx % w), cy = (int)(cur.idx / w);                            \
        for (int k = 0; k < 4; ++k) {                                                    \
            int px = cx + k_off_x[k], py = cy + k_off_y[k];                              \
            if (px < 0 || py < 0 || px >= (int)w || py >= (int)h) continue;              \
            size_t pi = (size_t)px + (size_t)py * w;                                      \
            if (!(cost[pi] < INF_T) || time_out[pi] != INF_T) continue;                 \
            KEY_T l = px > 0 ? time_out[pi - 1] : INF_T;                                 \
            KEY_T r = px + 1 < (int)w ? time_out[pi + 1] : INF_T;                        \
            KEY_T u = py > 0 ? time_out[pi - w] : INF_T;                                 \
            KEY_T d = py + 1 < (int)h ? time_out[pi + w] : INF_T;                        \
            KEY_T ax = l < r ? l : r, ay = u < d ? u : d;                                \
            KEY_T t = GODUNOV(ax, ay, cost[pi], INF_T);                                  \
            if (HEAP##_push(&hp, t, (uint32_t)pi)) { free(hp.a); return -1; }            \
        }                                                                                \
    }                                                                                    \
    free(hp.a);                                                                          \
    return 0;

int orc_fmm_f32(const float* cost, uint32_t w, uint32_t h,
                const uint32_t* seeds, uint32_t n_seeds, float* time_out) {
    FMM_BODY(float, heap32, orc_godunov_cpu, INFINITY)
}

int orc_fmm_f64(const double* cost, uint32_t w, uint32_t h,
                const uint32_t* seeds, uint32_t n_seeds, double* time_out) {
    FMM_BODY(double, heap64, orc_godunov_f64, (double)INFINITY)
}

/* ------------------------------------------------------------------------- */
/* FIM (src/eikonal_fim.cpp:13-107)                                            */
/* ------------------------------------------------------------------------- */

typedef struct { uint32_t* a; size_t cap, head, n; } ring_t;

static int ring_push(ring_t* q, uint32_t v) {
    if (q->n == q->cap) {
        size_t nc = q->cap ? q->cap * 2 : 1024;
        uint32_t* na = (uint32_t*)malloc(nc * sizeof(uint32_t));
        if (!na) return -1;
        for (size_t i = 0; i < q->n; ++i) na[i] = q->a[(q->head + i) % q->cap];
        free(q->a); q->a = na; q->cap = nc; q->head = 0;
    }
    q->a[(q->head + q->n) % q->cap] = v;
    q->n++;
    return 0;
}

static uint32_t ring_pop(ring_t* q) {
    uint32_t v = q->a[q->head];
    q->head = (q->head + 1) % q->cap;
    q->n--;
    return v;
}

typedef struct {
    const float* cost; float* t; uint32_t w, h;
} fim_ctx;

static float fim_get_t(const fim_ctx* c, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)c->w || y >= (int)c->h) return INFINITY;
    return c->t[(size_t)x + (size_t)y * c->w];
}
static float fim_get_c(const fim_ctx* c, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)c->w || y >= (int)c->h) return INFINITY;
    return c->cost[(size_t)x + (size_t)y * c->w];
}

/* eikonal_solve_helper (src/eikonal_fim.cpp:35-47): returns new time, sets *diff
 * with eikonal_time_diff semantics (src/eikonal.h:142-143). */
static float fim_helper(const fim_ctx* c, int x, int y, float* diff) {
    float ax = fminf(fim_get_t(c, x + 1, y), fim_get_t(c, x - 1, y));
    float ay = fminf(fim_get_t(c, x, y + 1), fim_get_t(c, x, y - 1));
    float told = fim_get_t(c, x, y);
    float tnew = orc_godunov_cpu(ax, ay, fim_get_c(c, x, y), told);
    *diff = (tnew == told) ? 0.0f : (tnew - told);
    return tnew;
}

int64_t orc_fim_f32(float minimum_change, const float* cost, uint32_t w, uint32_t h,
                    const uint32_t* seeds, uint32_t n_seeds, float* time_out) {
    if (!cost || !time_out || w == 0 || h == 0 || (n_seeds && !seeds)) return -1;
    size_t N = (size_t)w * h;
    for (size_t i = 0; i < N; ++i) time_out[i] = INFINITY;
    uint8_t* open = (uint8_t*)calloc(N, 1);
    ring_t q = { 0, 0, 0, 0 };
    if (!open) return -1;
    fim_ctx c = { cost, time_out, w, h };
    /* seed order of src/eikonal_fim.cpp:83-94: (-1,0),(0,-1),(1,0),(0,1) */
    static const int sx[4] = { -1, 0, 1, 0 }, sy[4] = { 0, -1, 0, 1 };
    for (uint32_t s = 0; s < n_seeds; ++s) {
        uint32_t x = seeds[2 * s], y = seeds[2 * s + 1];
        if (x >= w || y >= h) { free(open); free(q.a); return -1; }
        time_out[x + (size_t)y * w] = 0.0f;
        for (int k = 0; k < 4; ++k) {
            int px = (int)x + sx[k], py = (int)y + sy[k];
            if (px < 0 || py < 0 || px >= (int)w || py >= (int)h) continue;
            open[(size_t)px + (size_t)py * w] = 1;
            if (ring_push(&q, (uint32_t)px + (uint32_t)py * w)) { free(open); free(q.a); return -1; }
        }
    }
    int64_t solves = 0;
    while (q.n) {
        uint32_t p = ring_pop(&q);
        int x = (int)(p % w), y = (int)(p / w);
        float diff;
        float tnew = fim_helper(&c, x, y, &diff);
        ++solves;
        time_out[p] = tnew;                     /* unconditional, :46 */
        if (minimum_change < fabsf(diff)) {      /* :49-51 */
            if (ring_push(&q, p)) { free(open); free(q.a); return -1; }
            continue;
        }
        for (int k = 0; k < 4; ++k) {            /* x axis then y axis, -axis then +axis */
            int px = x + k_off_x[k], py = y + k_off_y[k];
            if (fim_get_c(&c, px, py) == INFINITY) continue;
            size_t pi = (size_t)px + (size_t)py * w;
            if (open[pi]) continue;
            float nd;
            float nt = fim_helper(&c, px, py, &nd);
            ++solves;
            if (nd < 0) {
                time_out[pi] = nt;
                open[pi] = 1;
                if (ring_push(&q, (uint32_t)pi)) { free(open); free(q.a); return -1; }
            }
        }
        open[p] = 0;
    }
    free(open); free(q.a);
    return solves;
}

/* ------------------------------------------------------------------------- */
/* Brute force: fast-sweeping Gauss-Seidel fixed point in FP64                 */
/* ------------------------------------------------------------------------- */

int orc_sweep_f64(const double* cost, uint32_t w, uint32_t h,
                  const uint32_t* seeds, uint32_t n_seeds, double* t, int max_rounds) {
    if (!cost || !t || w == 0 || h == 0) return -1;
    size_t N = (size_t)w * h;
    uint8_t* fixed = (uint8_t*)calloc(N, 1);
    if (!fixed) return -1;
    for (size_t i = 0; i < N; ++i) t[i] = INFINITY;
    for (uint32_t s = 0; s < n_seeds; ++s) {
        uint32_t x = seeds[2 * s], y = seeds[2 * s + 1];
        if (x >= w || y >= h) { free(fixed); return -1; }
        t[x + (size_t)y * w] = 0.0; fixed[x + (size_t)y * w] = 1;
    }
    int rounds = 0;
    for (; rounds < max_rounds; ++rounds) {
        int changed = 0;
        for (int dir = 0; dir < 4; ++dir) {
            int xs = (dir == 0 || dir == 3) ? 1 : -1;
            int ys = (dir < 2) ? 1 : -1;
            for (uint32_t jj = 0; jj < h; ++jj) {
                int y = ys > 0 ? (int)jj : (int)(h - 1 - jj);
                for (uint32_t ii = 0; ii < w; ++ii) {
                    int x = xs > 0 ? (int)ii : (int)(w - 1 - ii);
                    size_t i = (size_t)x + (size_t)y * w;
                    if (fixed[i]) continue;
                    double l = x > 0 ? t[i - 1] : INFINITY;
                    double r = x + 1 < (int)w ? t[i + 1] : INFINITY;
                    double u = y > 0 ? t[i - w] : INFINITY;
                    double d = y + 1 < (int)h ? t[i + w] : INFINITY;
                    double nv = orc_godunov_f64(l < r ? l : r, u < d ? u : d, cost[i], t[i]);
                    if (nv < t[i]) { t[i] = nv; changed = 1; }
                }
            }
        }
        if (!changed) { ++rounds; break; }
    }
    free(fixed);
    return rounds;
}

/* ------------------------------------------------------------------------- */
/* K-axis update and the three-dimensional solvers                             */
/* ------------------------------------------------------------------------- */

/* eikonal_solve<K> (src/eikonal.h:73-118): axis minima sorted ascending (the bubble sort
 * of eikonal_axis_sort, data/eikonal_common.cl:215-219), then for i = 0..K-1: stop once
 * the best time does not exceed axis value i; stop at an INF axis; otherwise accumulate
 * the quadratic sum_{j<=i} (T - a_j)^2 = cost^2 (b += -2 a_i, c += a_i^2 with c starting
 * at -cost^2) and take its root, a_0 + cost at i = 0.  The reference's release build
 * takes the smaller root (-b - sqrt(quad)) / 2a at :101, which lies below a_i; the
 * scheme's root is the larger one, taken here. */
#define GODUNOV_K_BODY(T, SQRT)                                                            \
    T v[3];                                                                                \
    if (K < 1 || K > 3) return best_prev;                                                   \
    for (int i = 0; i < K; ++i) v[i] = dims[i];                                             \
    for (int i = 0; i < K; ++i)                                                             \
        for (int j = 0; j + 1 < K - i; ++j)                                                 \
            if (v[j + 1] < v[j]) { T x = v[j]; v[j] = v[j + 1]; v[j + 1] = x; }             \
    T b = 0, c = -(cost * cost), best = best_prev;                                          \
    for (int i = 0; i < K; ++i) {                                                           \
        T ta = v[i];                                                                        \
        if (best <= ta) break;                                                              \
        if (isinf(ta)) return best;                                                         \
        b += -2 * ta;                                                                       \
        c += ta * ta;                                                                       \
        if (i == 0) {                                                                       \
            T t = ta + cost;                                                                \
            best = t < best ? t : best;                                                     \
        } else {                                                                            \
            T a = (T)(i + 1);                                                               \
            T quad = b * b - 4 * a * c;                                                     \
            if (quad < 0) quad = 0;                                                         \
            T t = (-b + SQRT(quad)) / (2 * a);                                              \
            best = t < best ? t : best;                                                     \
        }                                                                                   \
    }                                                                                       \
    return best;

float orc_godunov_k_f32(const float* dims, int K, float cost, float best_prev) {
    GODUNOV_K_BODY(float, sqrtf)
}

double orc_godunov_k_f64(const double* dims, int K, double cost, double best_prev) {
    GODUNOV_K_BODY(double, sqrt)
}

/* Neighbour visit order: -x, +x, -y, +y (src/eikonal_fmm.cpp:61-62), then -z, +z. */
static const int k3_off[6][3] = { {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1} };

/* Axis minima of cell (x, y, z) of a W x H x D field, out-of-grid reads +INF. */
#define AXIS_MINS_3D(T, t, x, y, z, out)                                                    \
    do {                                                                                    \
        size_t ci_ = (size_t)(x) + (size_t)w * ((size_t)(y) + (size_t)h * (size_t)(z));     \
        T l_ = (x) > 0 ? t[ci_ - 1] : (T)INFINITY;                                           \
        T r_ = (x) + 1 < (int)w ? t[ci_ + 1] : (T)INFINITY;                                  \
        T u_ = (y) > 0 ? t[ci_ - w] : (T)INFINITY;                                           \
        T d_ = (y) + 1 < (int)h ? t[ci_ + w] : (T)INFINITY;                                  \
        T f_ = (z) > 0 ? t[ci_ - (size_t)w * h] : (T)INFINITY;                               \
        T b_ = (z) + 1 < (int)dd ? t[ci_ + (size_t)w * h] : (T)INFINITY;                     \
        out[0] = l_ < r_ ? l_ : r_;                                                         \
        out[1] = u_ < d_ ? u_ : d_;                                                         \
        out[2] = f_ < b_ ? f_ : b_;                                                         \
    } while (0)

/* FMM in three dimensions: the sequential FMM of src/eikonal_fmm.cpp:12-71 with six
 * neighbours and the K = 3 update from accepted values. */
#define FMM3D_BODY(KEY_T, HEAP, GODUNOV_K)                                                   \
    if (!cost || !time_out || w == 0 || h == 0 || dd == 0 || (n_seeds && !seeds)) return -1; \
    size_t N = (size_t)w * h * dd;                                                           \
    for (size_t i = 0; i < N; ++i) time_out[i] = (KEY_T)INFINITY;                            \
    HEAP hp = { 0, 0, 0 };                                                                   \
    for (uint32_t s = 0; s < n_seeds; ++s) {                                                 \
        uint32_t x = seeds[3 * s], y = seeds[3 * s + 1], z = seeds[3 * s + 2];               \
        if (x >= w || y >= h || z >= dd) { free(hp.a); return -1; }                          \
        if (HEAP##_push(&hp, (KEY_T)0, (uint32_t)(x + (size_t)w * (y + (size_t)h * z)))) {   \
            free(hp.a); return -1;                                                           \
        }                                                                                    \
    }                                                                                        \
    while (hp.n) {                                                                           \
        HEAP##_ent cur = HEAP##_pop(&hp);                                                    \
        if (time_out[cur.idx] <= cur.key) continue;                                          \
        time_out[cur.idx] = cur.key;                                                         \
        int cx = (int)(cur.idx % w), cy = (int)((cur.idx / w) % h), cz = (int)(cur.idx / ((size_t)w * h)); \
        for (int k = 0; k < 6; ++k) {                                                        \
            int px = cx + k3_off[k][0], py = cy + k3_off[k][1], pz = cz + k3_off[k][2];      \
            if (px < 0 || py < 0 || pz < 0 || px >= (int)w || py >= (int)h || pz >= (int)dd) continue; \
            size_t pi = (size_t)px + (size_t)w * ((size_t)py + (size_t)h * (size_t)pz);      \
            if (!(cost[pi] < (KEY_T)INFINITY) || time_out[pi] != (KEY_T)INFINITY) continue;  \
            KEY_T m[3];                                                                      \
            AXIS_MINS_3D(KEY_T, time_out, px, py, pz, m);                                    \
            KEY_T t = GODUNOV_K(m, 3, cost[pi], (KEY_T)INFINITY);                            \
            if (HEAP##_push(&hp, t, (uint32_t)pi)) { free(hp.a); return -1; }                \
        }                                                                                    \
    }                                                                                        \
    free(hp.a);                                                                              \
    return 0;

int orc_fmm3d_f32(const float* cost, uint32_t w, uint32_t h, uint32_t dd,
                  const uint32_t* seeds, uint32_t n_seeds, float* time_out) {
    FMM3D_BODY(float, heap32, orc_godunov_k_f32)
}

int orc_fmm3d_f64(const double* cost, uint32_t w, uint32_t h, uint32_t dd,
                  const uint32_t* seeds, uint32_t n_seeds, double* time_out) {
    FMM3D_BODY(double, heap64, orc_godunov_k_f64)
}

/* Fast-sweeping Gauss-Seidel in three dimensions (eight orderings) to a fixed point, FP64. */
int orc_sweep3d_f64(const double* cost, uint32_t w, uint32_t h, uint32_t dd,
                    const uint32_t* seeds, uint32_t n_seeds, double* t, int max_rounds) {
    if (!cost || !t || w == 0 || h == 0 || dd == 0) return -1;
    size_t N = (size_t)w * h * dd;
    uint8_t* fixed = (uint8_t*)calloc(N, 1);
    if (!fixed) return -1;
    for (size_t i = 0; i < N; ++i) t[i] = INFINITY;
    for (uint32_t s = 0; s < n_seeds; ++s) {
        uint32_t x = seeds[3 * s], y = seeds[3 * s + 1], z = seeds[3 * s + 2];
        if (x >= w || y >= h || z >= dd) { free(fixed); return -1; }
        size_t i = x + (size_t)w * (y + (size_t)h * z);
        t[i] = 0.0; fixed[i] = 1;
    }
    int rounds = 0;
    for (; rounds < max_rounds; ++rounds) {
        int changed = 0;
        for (int dir = 0; dir < 8; ++dir) {
            int xs = (dir & 1) ? -1 : 1, ys = (dir & 2) ? -1 : 1, zs = (dir & 4) ? -1 : 1;
            for (uint32_t kk = 0; kk < dd; ++kk) {
                int z = zs > 0 ? (int)kk : (int)(dd - 1 - kk);
                for (uint32_t jj = 0; jj < h; ++jj) {
                    int y = ys > 0 ? (int)jj : (int)(h - 1 - jj);
                    for (uint32_t ii = 0; ii < w; ++ii) {
                        int x = xs > 0 ? (int)ii : (int)(w - 1 - ii);
                        size_t i = (size_t)x + (size_t)w * ((size_t)y + (size_t)h * (size_t)z);
                        if (fixed[i]) continue;
                        double m[3];
                        AXIS_MINS_3D(double, t, x, y, z, m);
                        double nv = orc_godunov_k_f64(m, 3, cost[i], t[i]);
                        if (nv < t[i]) { t[i] = nv; changed = 1; }
                    }
                }
            }
        }
        if (!changed) { ++rounds; break; }
    }
    free(fixed);
    return rounds;
}

/* ------------------------------------------------------------------------- */
/* Error metrics                                                                */
/* ------------------------------------------------------------------------- */

void orc_error_metrics(const float* got, const double* ref, size_t n, double out[5]) {
    double linf = 0.0, l1 = 0.0, maxref = 0.0, mism = 0.0, cnt = 0.0;
    for (size_t i = 0; i < n; ++i) {
        int fg = isfinite(got[i]), fr = isfinite(ref[i]);
        if (fr && ref[i] > maxref) maxref = ref[i];
        if (fg != fr) { mism += 1.0; continue; }
        if (!fg) continue;
        double e = fabs((double)got[i] - ref[i]);
        if (e > linf) linf = e;
        l1 += e;
        cnt += 1.0;
    }
    out[0] = linf;
    out[1] = cnt > 0 ? l1 / cnt : 0.0;
    out[2] = maxref;
    out[3] = mism;
    out[4] = cnt;
}
