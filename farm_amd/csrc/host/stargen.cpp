// Seeded synthetic star-shaped polygon sets, generated in parallel on the
// host: the workloads' polygon generators (star pairs, complex polygons,
// lakes) at full size -- hundreds of millions of vertices in about a second.
//
// Polygon i draws from its own counter-based stream (splitmix64 of the seed
// and i), so a set does not depend on the thread count.  Per polygon: the
// centre uniform in [lo, hi]^2, the vertex count uniform in [nv_lo, nv_hi],
// R ~ LogNormal(r_mu, r_sigma) (Box-Muller), sorted angles as normalised
// partial sums of nv + 1 exponentials (the order statistics of nv uniforms on
// [0, 2 pi)), radius R (1 + amp u) with u ~ U(-1, 1) per vertex; coordinates
// rounded to float32; rings closed (the first vertex repeated).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../../include/farm_host.h"
#include "raster.h"

namespace {

inline uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t s;
    Stream(uint64_t seed, uint64_t i, uint64_t which) {
        s = seed * 0x632be59bd9b4e019ull + i * 0x9e3779b97f4a7c15ull + which * 0xd1b54a32d192ed03ull;
        splitmix(s);
    }
    double uniform() {  // (0, 1)
        return ((splitmix(s) >> 11) + 0.5) * 0x1p-53;
    }
    double normal() {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
    }
};

}  // namespace

namespace {

// Shared body: vertex counts uniform in [nv_lo, nv_hi], or with
// pareto_alpha > 0 Pareto(alpha) with scale nv_lo, floored and clipped to
// [nv_lo, nv_hi]; noise 0: r = R (1 + amp u), u ~ U(-1, 1) per vertex;
// noise 1: three octaves, r = R (1 + amp s / 1.75),
// s = sum_o 0.5^o sin(f 2^o theta + phi_o), f ~ U{2..6}, phi_o ~ U[0, 2 pi)
// per polygon.
farm_hostset *star_set(uint64_t seed, int64_t n, double lo, double hi, int nv_lo, int nv_hi, double pareto_alpha,
                       double r_mu, double r_sigma, int noise, double amp, int nthreads) {
    if (n < 0 || nv_lo < 3 || nv_hi < nv_lo || !(hi >= lo) || noise < 0 || noise > 1 || pareto_alpha < 0)
        return nullptr;
    const int nth = std::max(1, nthreads);
    // vertex counts first (their own streams), then the ring offsets
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        Stream r(seed, (uint64_t)i, 0);
        int nv;
        if (pareto_alpha > 0) {
            const double v = std::floor(nv_lo * std::pow(r.uniform(), -1.0 / pareto_alpha));
            nv = (int)std::min<double>(nv_hi, std::max<double>(nv_lo, v));
        } else {
            nv = nv_lo + (int)std::min<double>(nv_hi - nv_lo, std::floor(r.uniform() * (nv_hi - nv_lo + 1)));
        }
        off[i + 1] = off[i] + (uint64_t)nv + 1;
    }
    std::vector<float> xy(2 * off[n]);
    auto work = [&](int t) {
        std::vector<double> e;
        for (int64_t i = t; i < n; i += nth) {
            Stream r(seed, (uint64_t)i, 1);
            const int nv = (int)(off[i + 1] - off[i]) - 1;
            const double cx = lo + (hi - lo) * r.uniform();
            const double cy = lo + (hi - lo) * r.uniform();
            const double R = std::exp(r_mu + r_sigma * r.normal());
            double f = 0, ph[3] = {0, 0, 0};
            if (noise == 1) {
                f = 2 + std::min(4.0, std::floor(r.uniform() * 5));
                for (double &p : ph) p = 2.0 * M_PI * r.uniform();
            }
            e.resize((size_t)nv + 1);
            double acc = 0;
            for (int k = 0; k <= nv; k++) {
                acc += -std::log(r.uniform());
                e[k] = acc;
            }
            const double inv = 2.0 * M_PI / acc;
            float *p = xy.data() + 2 * off[i];
            for (int k = 0; k < nv; k++) {
                const double th = e[k] * inv;
                double rad;
                if (noise == 1) {
                    double sn = 0;
                    for (int o = 0; o < 3; o++) sn += std::ldexp(1.0, -o) * std::sin(f * (1 << o) * th + ph[o]);
                    rad = R * (1.0 + amp * sn / 1.75);
                } else {
                    rad = R * (1.0 + amp * (2.0 * r.uniform() - 1.0));
                }
                p[2 * k] = (float)(cx + rad * std::cos(th));
                p[2 * k + 1] = (float)(cy + rad * std::sin(th));
            }
            p[2 * nv] = p[0];
            p[2 * nv + 1] = p[1];
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return farm::hostset_adopt(std::move(off), std::move(xy), nth);
}

}  // namespace

extern "C" farm_hostset *farm_hostset_star_polygons(uint64_t seed, int64_t n, double lo, double hi, int nv_lo,
                                                    int nv_hi, double r_mu, double r_sigma, double amp,
                                                    int nthreads) {
    return star_set(seed, n, lo, hi, nv_lo, nv_hi, 0.0, r_mu, r_sigma, 0, amp, nthreads);
}

extern "C" farm_hostset *farm_hostset_star_polygons_ex(uint64_t seed, int64_t n, double lo, double hi, int nv_lo,
                                                       int nv_hi, double pareto_alpha, double r_mu, double r_sigma,
                                                       int noise, double amp, int nthreads) {
    return star_set(seed, n, lo, hi, nv_lo, nv_hi, pareto_alpha, r_mu, r_sigma, noise, amp, nthreads);
}
