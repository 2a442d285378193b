// Seeded synthetic star-shaped polygon sets, generated in parallel on the
// host: the workloads' polygon generators at the sizes the learned part's
// training and evaluation need (millions of polygons per set).
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

extern "C" farm_hostset *farm_hostset_star_polygons(uint64_t seed, int64_t n, double lo, double hi, int nv_lo,
                                                    int nv_hi, double r_mu, double r_sigma, double amp,
                                                    int nthreads) {
    if (n < 0 || nv_lo < 3 || nv_hi < nv_lo || !(hi >= lo)) return nullptr;
    const int nth = std::max(1, nthreads);
    // vertex counts first (their own streams), then the ring offsets
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        Stream r(seed, (uint64_t)i, 0);
        const int nv = nv_lo + (int)std::min<double>(nv_hi - nv_lo, std::floor(r.uniform() * (nv_hi - nv_lo + 1)));
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
                const double rad = R * (1.0 + amp * (2.0 * r.uniform() - 1.0));
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
    return farm_hostset_from_rings(n, off.data(), xy.data());
}
