// field_gen.cpp -- seeded synthetic speed fields for the solver's benchmark workloads.
//
// The reference harness builds its cost fields from PNGs and closed forms
// (src/test_runner.cpp:32-76, 304-316) and converts speed to cost as 1/speed
// (:32-36).  The workloads here are seeded generators instead, defined exactly so
// any caller can reproduce them:
//
//   PRNG      splitmix64 (Steele et al. 2014): s += 0x9E3779B97F4A7C15;
//             z = (s ^ s>>30) * 0xBF58476D1CE4E5B9; z = (z ^ z>>27) * 0x94D049BB133111EB;
//             out = z ^ z>>31.  Uniform u = (out >> 11) * 2^-53 in [0, 1).
//   normal    Box-Muller on (u1, u2): r = sqrt(-2 ln(1 - u1)), z0 = r cos(2 pi u2),
//             z1 = r sin(2 pi u2); both used, row-major order.
//   rings     7 circular walls centred at (W/2, H/2) (integer division), radii
//             r_k = scale * (16 + k (120 - 16) / 6), k = 0..6; a cell (x, y) is wall when
//             | |(x, y) - c| - r_k | < 1.5 scale, except inside the gap: angular distance
//             to theta_k < 3 scale / r_k (a 6*scale-cell arc), theta_k = 2 pi u_k drawn in
//             order k = 0..6.  Wall speed `wall_speed`, background 1.
//   smooth    z iid N(0,1), Gaussian filter sigma (separable, radius int(4 sigma + 0.5),
//             weights exp(-i^2 / 2 sigma^2) normalised, half-sample 'reflect' boundary),
//             then log10 speed = log10 lo + (log10 hi - log10 lo)(z - min z)/(max z - min z).
//   points    x = floor(u W), y = floor(u H), drawn as (x0, y0, x1, y1, ...).
#include "capi_common.h"
#include "ghcm.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

int reflect_index(long i, long n) {
    const long period = 2 * n;
    long m = i % period;
    if (m < 0) m += period;
    return (int)(m < n ? m : period - 1 - m);
}

// 1-D Gaussian pass along rows (stride 1) or columns (stride w) over `count` lines.
void gauss_pass(const double* in, double* out, int w, int h, bool along_x,
                const std::vector<double>& wt, int radius) {
    const int lines = along_x ? h : w;
    const int len = along_x ? w : h;
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int l0, int l1) {
        std::vector<double> buf(len);
        std::vector<int> ref_idx;
        for (int l = l0; l < l1; ++l) {
            for (int i = 0; i < len; ++i)
                buf[i] = along_x ? in[(size_t)l * w + i] : in[(size_t)i * w + l];
            for (int i = 0; i < len; ++i) {
                double acc = 0.0;
                if (i - radius >= 0 && i + radius < len) {
                    const double* b = &buf[i - radius];
                    for (int k = 0; k <= 2 * radius; ++k) acc += wt[k] * b[k];
                } else {
                    for (int k = -radius; k <= radius; ++k)
                        acc += wt[k + radius] * buf[reflect_index((long)i + k, len)];
                }
                if (along_x) out[(size_t)l * w + i] = acc;
                else out[(size_t)i * w + l] = acc;
            }
        }
    };
    std::vector<std::thread> th;
    int per = (lines + (int)nthreads - 1) / (int)nthreads;
    for (unsigned t = 0; t < nthreads; ++t) {
        int l0 = (int)t * per, l1 = std::min(lines, l0 + per);
        if (l0 >= l1) break;
        th.emplace_back(work, l0, l1);
    }
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" int ghcm_gen_ring_speed(uint32_t w, uint32_t h, uint64_t seed, double wall_speed,
                                   double scale, double* out) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!out || w == 0 || h == 0 || !(scale > 0.0) || !(wall_speed >= 0.0)) return GHCM_ERR_INVALID_ARG;
        SplitMix64 rng(seed);
        const double two_pi = 6.283185307179586476925;
        double theta[7], radius[7];
        for (int k = 0; k < 7; ++k) {
            theta[k] = two_pi * rng.uniform();
            radius[k] = scale * (16.0 + k * (120.0 - 16.0) / 6.0);
        }
        const double cx = (double)(w / 2), cy = (double)(h / 2);
        const double half_thick = 1.5 * scale;
        for (uint32_t y = 0; y < h; ++y) {
            for (uint32_t x = 0; x < w; ++x) {
                const double dx = (double)x - cx, dy = (double)y - cy;
                const double r = std::sqrt(dx * dx + dy * dy);
                double v = 1.0;
                for (int k = 0; k < 7; ++k) {
                    if (std::fabs(r - radius[k]) < half_thick) {
                        double ang = std::atan2(dy, dx);
                        if (ang < 0) ang += two_pi;
                        double da = std::fabs(ang - theta[k]);
                        da = std::min(da, two_pi - da);
                        if (!(da < 3.0 * scale / radius[k])) v = wall_speed;
                        break;
                    }
                }
                out[(size_t)y * w + x] = v;
            }
        }
        return GHCM_OK;
    });
}

extern "C" int ghcm_gen_smooth_speed(uint32_t w, uint32_t h, uint64_t seed, double sigma,
                                     double lo, double hi, double* out) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!out || w == 0 || h == 0 || !(sigma > 0.0) || !(lo > 0.0) || !(hi >= lo)) return GHCM_ERR_INVALID_ARG;
        const size_t n = (size_t)w * h;
        std::vector<double> z(n), tmp(n);
        SplitMix64 rng(seed);
        const double two_pi = 6.283185307179586476925;
        for (size_t i = 0; i < n; i += 2) {
            double u1 = rng.uniform(), u2 = rng.uniform();
            double r = std::sqrt(-2.0 * std::log(1.0 - u1));
            z[i] = r * std::cos(two_pi * u2);
            if (i + 1 < n) z[i + 1] = r * std::sin(two_pi * u2);
        }
        const int radius = (int)(4.0 * sigma + 0.5);
        std::vector<double> wt(2 * radius + 1);
        double sum = 0.0;
        for (int k = -radius; k <= radius; ++k) {
            wt[k + radius] = std::exp(-0.5 * (double)k * k / (sigma * sigma));
            sum += wt[k + radius];
        }
        for (double& v : wt) v /= sum;
        gauss_pass(z.data(), tmp.data(), (int)w, (int)h, true, wt, radius);
        gauss_pass(tmp.data(), z.data(), (int)w, (int)h, false, wt, radius);
        double zmin = z[0], zmax = z[0];
        for (double v : z) { zmin = std::min(zmin, v); zmax = std::max(zmax, v); }
        const double llo = std::log10(lo), lhi = std::log10(hi);
        const double span = zmax > zmin ? zmax - zmin : 1.0;
        for (size_t i = 0; i < n; ++i) out[i] = std::pow(10.0, llo + (lhi - llo) * (z[i] - zmin) / span);
        return GHCM_OK;
    });
}

extern "C" int ghcm_gen_points(uint32_t w, uint32_t h, uint64_t seed, uint32_t n, uint32_t* xy) {
    if ((!xy && n) || w == 0 || h == 0) return GHCM_ERR_INVALID_ARG;
    SplitMix64 rng(seed);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t x = (uint32_t)(rng.uniform() * w), y = (uint32_t)(rng.uniform() * h);
        xy[2 * i] = std::min(x, w - 1);
        xy[2 * i + 1] = std::min(y, h - 1);
    }
    return GHCM_OK;
}
