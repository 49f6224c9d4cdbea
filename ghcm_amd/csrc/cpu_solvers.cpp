// cpu_solvers.cpp -- the two sequential CPU solvers of the C++ API.
//
// eikonal_fmm_sequential follows src/eikonal_fmm.cpp:12-71: a min-heap with lazy
// deletion; a cell is accepted at its first pop and its 4 neighbours are (re)solved
// from accepted values only.  eikonal_fim_sequential follows src/eikonal_fim.cpp:13-107:
// a FIFO active list; a cell is re-queued while it changes by more than
// `minimum_change`, otherwise it opens each neighbour whose time strictly decreases.
// Both use the reference CPU Godunov form eikonal_solve<2> (eikonal.h), so results
// match the reference's CPU solvers up to heap tie order.
#include <chrono>
#include <deque>
#include <queue>

#include "eikonal.h"

namespace ghcm {

namespace {

struct Pending {
    float t;
    uint32_t idx;
    bool operator<(const Pending& o) const { return t > o.t; }   // min-heap
};

using clock_type = std::chrono::high_resolution_clock;

}  // namespace

EikonalResult eikonal_fmm_sequential(Field2D<float> const& cost, std::vector<uint2> const& seeds) {
    EikonalResult res;
    const uint2 sz = cost.size();
    Field2D<float> T(sz, INFINITY);
    const auto t_at = [&](int x, int y) { return T.get_or(INFINITY, int2(x, y)); };

    std::priority_queue<Pending> heap;
    const auto t0 = clock_type::now();
    for (const uint2& s : seeds) {
        if (!cost.in_bounds(s)) { res.status = GHCM_ERR_INVALID_ARG; return res; }
        heap.push({0.0f, (uint32_t)dim_idx(sz, s)});
    }
    static const int ox[4] = {-1, 1, 0, 0}, oy[4] = {0, 0, -1, 1};
    while (!heap.empty()) {
        const Pending cur = heap.top();
        heap.pop();
        if (T[cur.idx] <= cur.t) continue;          // accepted earlier with a smaller time
        T[cur.idx] = cur.t;
        const int cx = (int)(cur.idx % sz.x), cy = (int)(cur.idx / sz.x);
        for (int k = 0; k < 4; ++k) {
            const int px = cx + ox[k], py = cy + oy[k];
            if (!(cost.get_or(INFINITY, int2(px, py)) < INFINITY)) continue;
            if (t_at(px, py) != INFINITY) continue;
            float ax[2] = {std::min(t_at(px + 1, py), t_at(px - 1, py)),
                           std::min(t_at(px, py + 1), t_at(px, py - 1))};
            if (ax[1] < ax[0]) std::swap(ax[0], ax[1]);
            const float cc = cost[uint2((uint32_t)px, (uint32_t)py)];
            heap.push({eikonal_solve<2>(ax, cc, INFINITY), (uint32_t)(px + (size_t)py * sz.x)});
        }
    }
    const auto t1 = clock_type::now();
    res.duration_total = res.duration_device = std::chrono::duration_cast<EikonalResult::duration_t>(t1 - t0);
    res.field = std::move(T);
    return res;
}

EikonalResult eikonal_fim_sequential(float minimum_change, Field2D<float> const& cost,
                                     std::vector<uint2> const& seeds) {
    EikonalResult res;
    const uint2 sz = cost.size();
    Field2D<float> T(sz, INFINITY);
    std::vector<uint8_t> open(cost.size_flat(), 0);
    std::deque<uint32_t> active;

    const auto t_at = [&](int x, int y) { return T.get_or(INFINITY, int2(x, y)); };
    const auto c_at = [&](int x, int y) { return cost.get_or(INFINITY, int2(x, y)); };
    // New time and eikonal_time_diff(new, old) at (x, y).
    const auto update = [&](int x, int y, float* diff) {
        float ax[2] = {std::min(t_at(x + 1, y), t_at(x - 1, y)), std::min(t_at(x, y + 1), t_at(x, y - 1))};
        if (ax[1] < ax[0]) std::swap(ax[0], ax[1]);
        const float old = t_at(x, y);
        const float nv = eikonal_solve<2>(ax, c_at(x, y), old);
        *diff = eikonal_time_diff(nv, old);
        return nv;
    };

    const auto t0 = clock_type::now();
    static const int sx[4] = {-1, 0, 1, 0}, sy[4] = {0, -1, 0, 1};   // seed neighbour order
    for (const uint2& s : seeds) {
        if (!cost.in_bounds(s)) { res.status = GHCM_ERR_INVALID_ARG; return res; }
        T[s] = 0.0f;
        for (int k = 0; k < 4; ++k) {
            const int2 p(int2(s) + int2(sx[k], sy[k]));
            if (!cost.in_bounds(p)) continue;
            const uint32_t i = (uint32_t)(p.x + (size_t)p.y * sz.x);
            open[i] = 1;
            active.push_back(i);
        }
    }
    static const int nx[4] = {-1, 1, 0, 0}, ny[4] = {0, 0, -1, 1};
    while (!active.empty()) {
        const uint32_t i = active.front();
        active.pop_front();
        const int x = (int)(i % sz.x), y = (int)(i / sz.x);
        float diff;
        T[i] = update(x, y, &diff);
        if (minimum_change < std::fabs(diff)) {     // not converged: solve again later
            active.push_back(i);
            continue;
        }
        for (int k = 0; k < 4; ++k) {
            const int px = x + nx[k], py = y + ny[k];
            if (c_at(px, py) == INFINITY) continue;
            const uint32_t j = (uint32_t)(px + (size_t)py * sz.x);
            if (open[j]) continue;
            float d;
            const float nv = update(px, py, &d);
            if (d < 0) {
                T[j] = nv;
                open[j] = 1;
                active.push_back(j);
            }
        }
        open[i] = 0;
    }
    const auto t1 = clock_type::now();
    res.duration_total = res.duration_device = std::chrono::duration_cast<EikonalResult::duration_t>(t1 - t0);
    res.field = std::move(T);
    return res;
}

}  // namespace ghcm
