// Within-distance decision of candidate pairs on gfx950.
//
// Follows the CPU semantics of Ideal::within (src/geometry/Ideal.cpp:2475-2643)
// with the distance helpers point_to_segment_distance /
// segment_to_segment_within_batch (src/include/geometry_computation.h:49-182)
// and box::distance / box::max_distance (src/geometry/Box.cpp:89-149).  The
// reference GPU path (src/cuda/query_within_polygon.cu) computes the
// refinement in FP32 and decides differently; this one reproduces the CPU
// decision bit for bit.
//
// One wavefront per candidate pair; all control flow of a pair is uniform
// across the wave and the lanes split the wide loops:
//  1. hierarchical descent: the pixel sets of both polygons are kept as
//     compact lists of packed (x, y) ids in LDS.  The reference's queue of
//     pixel pairs only feeds the union of its first members and the union of
//     its second members into the next level, so the two projections carry
//     the same information; at the last level the candidate pixel pairs are
//     the filtered cross product of the two lists.
//  2. box distances: box::max_distance over the 16 corner pairs reduces to
//     the largest |dx| and |dy| (the corner sets are a full product and every
//     step after the f32 differences is monotone).  Distances are compared
//     as squares against the squared bounds' thresholds (sq_le / sq_lt: the
//     correctly rounded root is monotone), so a square root is taken only
//     for the smallest maximum of a level and for kept candidates.  The
//     divisions by the km table entries, the latitude constant and the
//     raster steps go through precomputed reciprocals with two exact fma
//     corrections (div_rcp), which give the same rounded quotient.
//  3. candidates (apx, low, high, id1, id2) of the last level are generated
//     in passes over the last level's pixel pairs: each pass keeps the 64
//     smallest tuples after a bound (staged in LDS, merged into a sorted set
//     of 64 by bitonic networks, candidate k on lane k) and only the low
//     bound of the rest; the refinement takes the 64 in tuple order (the
//     order of std::sort over the tuples) with each one's reference_dist
//     from a suffix minimum, and a further pass follows only when all 64 are
//     refined without a decision.  Nothing is written to global memory.  (A
//     pair whose layers exceed the LDS lists runs on the global path, with
//     its candidates in a per-wave slab, taken by argmin selection.)
//  4. the truncated point-segment distances of one edge-sequence pair are
//     spread over the lanes; the minimum is taken on the squared distances
//     and one square root follows (sqrt is monotone and correctly rounded).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <type_traits>
#include <vector>

#include "handle.h"
#include "wave.h"

namespace farm {

const double *km_lon_table();  // host table shared with candidate generation

__constant__ double c_km_lon[900];  // degree_per_kilometer_longitude (util.h:73-96)
__constant__ double c_km_lon_inv[900];  // 1 / each entry, rounded to nearest (host division)
__constant__ float c_mean[20];      // mean[20] (util.h:68)

namespace {

constexpr double kKmLat = 360.0 / 40076.0;  // degree_per_kilometer_latitude (util.h:72)
constexpr double kKmLatInv = 1.0 / kKmLat;  // rounded to nearest (constant folding is IEEE)
constexpr double kFar = 100000.0;           // the reference's "no distance yet"
constexpr int kWaves = 4;                   // wavefronts per workgroup
constexpr int kPixMax = 512;                // pixels of a layer handled in LDS
constexpr int kBitWords = kPixMax / 32;
constexpr uint32_t kChunk = 8;              // pairs per queue grab

struct Cand {
    double apx, lo, hi;
    uint32_t id1, id2;
};

struct LGrid {
    double lx, ly, step;
    int dimx, dimy;
    const uint8_t *st;
};

__device__ __forceinline__ LGrid lgrid(const DevSet &S, const farm_raster_hdr &h, int k) {
    LGrid g;
    if (k == h.nlayers - 1) {
        g.lx = h.mbr[0];
        g.ly = h.mbr[1];
        g.step = h.step[0];
        g.dimx = h.dimx;
        g.dimy = h.dimy;
        g.st = S.status + h.pix_start;
    } else {
        const farm_layer &L = S.layers[h.layer_start + k];
        g.lx = L.low[0];
        g.ly = L.low[1];
        g.step = L.step;
        g.dimx = L.dimx;
        g.dimy = L.dimy;
        g.st = S.lstatus + L.status_start;
    }
    return g;
}

__device__ __forceinline__ bool border(uint8_t s) { return s != 0 && s != kCC - 1; }

// degree_per_kilometer_longitude(lat) (util.h:88-96); the index is clamped
// so that |lat| > 90 cannot read outside the table
__device__ __forceinline__ int km_idx(double lat) {
    double a = fabs(lat);
    if (a == 90) a = 89.9;
    const int i = (int)(a * 10);
    return min(max(i, 0), 899);
}

// The table entries one pair can need (its latitude span is a few tenths of
// a degree), copied per wave into LDS: a lookup is an LDS read instead of a
// dependent global load; an index outside the copy reads the table.
constexpr int kKmC = 16;
struct KmC {
    double v[kKmC], inv[kKmC];
    int lo, n;
};

// a / b rounded to nearest, for b > 0 (a table entry, the latitude
// constant or a raster step, 2^-20 .. 2^10) and y = RN(1 / b): q = a y, then two corrections
// q' = q + (a - b q) y by fma.  The residual a - b q of a q within one ulp of
// a / b is exact, and with y within half an ulp of 1 / b one correction of
// such a q is the rounded quotient (Markstein); the first brings a y (within
// 1.5 ulp) within one ulp, the second rounds.  A dividend whose residual
// could leave the normal range (|a| < 2^-900, not 0) or that is not finite
// takes the division.
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
    double q = a * y;
    q = fma(fma(-b, q, a), y, q);
    q = fma(fma(-b, q, a), y, q);
    const double m = fabs(a);
    if (__builtin_expect(!(m < 0x1p1000) || (m < 0x1p-900 && m != 0.0), 0)) q = a / b;
    return q;
}

// x / degree_per_kilometer_longitude(lat)
__device__ __forceinline__ double div_km_lon(double x, double lat, const KmC &c) {
    const int i = km_idx(lat);
    const int k = i - c.lo;
    const bool in = (unsigned)k < (unsigned)c.n;
    const double b = in ? c.v[k] : c_km_lon[i];
    const double y = in ? c.inv[k] : c_km_lon_inv[i];
    return div_rcp(x, b, y);
}

__device__ __forceinline__ double div_km_lat(double x) { return div_rcp(x, kKmLat, kKmLatInv); }

// MyRaster::get_pixel_box (MyRaster.cpp:312-322), square pixels
__device__ __forceinline__ void pbox(const LGrid &g, int x, int y, double b[4]) {
    b[0] = g.lx + x * g.step;
    b[1] = g.ly + y * g.step;
    b[2] = g.lx + (x + 1) * g.step;
    b[3] = g.ly + (y + 1) * g.step;
}

// box::distance(box, geography) (Box.cpp:89-120), b = this, before the
// square root (the root is monotone and correctly rounded: comparisons of
// the distance against a bound are comparisons of this value against the
// bound's threshold, sq_le / sq_lt)
__device__ __forceinline__ double box_min_d2(const double b[4], const double t[4], const KmC &kc) {
    if (!(t[0] > b[2] || t[2] < b[0] || t[1] > b[3] || t[3] < b[1])) return 0.0;
    double dx = 0, dy = 0;
    if (t[1] > b[3]) dy = t[1] - b[3];
    else if (t[3] < b[1]) dy = b[1] - t[3];
    if (t[0] > b[2]) dx = t[0] - b[2];
    else if (t[2] < b[0]) dx = b[0] - t[2];
    dy = div_km_lat(dy);
    dx = div_km_lon(dx, b[1], kc);
    return dx * dx + dy * dy;
}

// box::max_distance(box, geography) (Box.cpp:122-149): corners as f32
// points, f32 differences promoted to double; the largest |difference| per
// axis gives the maximum over the 16 corner pairs.  Before the square root.
__device__ __forceinline__ double box_max_d2(const double b[4], const double t[4], const KmC &kc) {
    const float b0 = (float)b[0], b1 = (float)b[1], b2 = (float)b[2], b3 = (float)b[3];
    const float t0 = (float)t[0], t1 = (float)t[1], t2 = (float)t[2], t3 = (float)t[3];
    const float fx = fmaxf(fmaxf(fabsf(b0 - t0), fabsf(b0 - t2)), fmaxf(fabsf(b2 - t0), fabsf(b2 - t2)));
    const float fy = fmaxf(fmaxf(fabsf(b1 - t1), fabsf(b1 - t3)), fmaxf(fabsf(b3 - t1), fabsf(b3 - t3)));
    const double dy = div_km_lat((double)fy);
    const double dx = div_km_lon((double)fx, b[1], kc);
    return dx * dx + dy * dy;
}

// The largest v >= 0 with sqrt(v) <= d (sq_le) or sqrt(v) < d (sq_lt), -1
// when there is none: sqrt(v) <= d <=> v <= sq_le(d), as the correctly
// rounded square root is monotone.
__device__ __forceinline__ double sq_le(double d) {
    if (!(d >= 0)) return -1.0;
    if (isinf(d)) return INFINITY;
    double v = d * d;
    while (v > 0 && sqrt(v) > d) v = nextafter(v, 0.0);
    while (sqrt(nextafter(v, INFINITY)) <= d) v = nextafter(v, INFINITY);
    return v;
}
__device__ __forceinline__ double sq_lt(double d) {
    if (!(d > 0)) return -1.0;
    if (isinf(d)) return INFINITY;
    double v = d * d;
    while (v > 0 && !(sqrt(v) < d)) v = nextafter(v, 0.0);
    if (!(sqrt(v) < d)) return -1.0;
    while (sqrt(nextafter(v, INFINITY)) < d) v = nextafter(v, INFINITY);
    return v;
}

// point_to_segment_distance (geometry_computation.h:49-89), geography on,
// before the square root
__device__ __forceinline__ double pt_seg_d2(float px, float py, float ax, float ay, float bx, float by,
                                            const KmC &kc) {
    const double A = (double)(px - ax);
    const double B = (double)(py - ay);
    const double C = (double)(bx - ax);
    const double D = (double)(by - ay);
    const double dot = A * C + B * D;
    const double len_sq = C * C + D * D;
    double param = -1;
    if (len_sq != 0) param = dot / len_sq;
    double xx, yy;
    if (param < 0) {
        xx = ax;
        yy = ay;
    } else if (param > 1) {
        xx = bx;
        yy = by;
    } else {
        xx = ax + param * C;
        yy = ay + param * D;
    }
    double dx = px - xx;
    double dy = py - yy;
    dx = div_km_lon(dx, py, kc);
    dy = div_km_lat(dy);
    return dx * dx + dy * dy;
}

// classifyPixel(area, pixelArea, count) (geometry_computation.h:299-321)
__device__ __forceinline__ int classify(double area, double pixel_area, int count) {
    const double ratio = area / pixel_area;
    if (fabs(ratio - 1.0) < 1e-9) return count - 1;
    if (fabs(ratio) < 1e-9) return 0;
    int idx = (int)ceil(ratio * (count - 2));
    if (idx >= count) idx = count - 1;
    return idx;
}





__device__ __forceinline__ bool tuple_less(const Cand &a, const Cand &b) {
    if (a.apx != b.apx) return a.apx < b.apx;
    if (a.lo != b.lo) return a.lo < b.lo;
    if (a.hi != b.hi) return a.hi < b.hi;
    if (a.id1 != b.id1) return a.id1 < b.id1;
    return a.id2 < b.id2;
}

// Pixel lists hold packed (x | y << SH): 16-bit entries with SH = 8 in the
// LDS path (layers up to 256 x 256), 32-bit entries with SH = 16 otherwise.
template <int SH, class T>
__device__ __forceinline__ T pack_xy(uint32_t x, uint32_t y) {
    return (T)(x | (y << SH));
}
template <int SH, class T>
__device__ __forceinline__ int unpack_x(T v) {
    return (int)((uint32_t)v & ((1u << SH) - 1u));
}
template <int SH, class T>
__device__ __forceinline__ int unpack_y(T v) {
    return (int)((uint32_t)v >> SH);
}

// Pixel set -> compact list (ascending ids): rounds of 64 ids, one per lane,
// each member placed by the ballot count of the members below it.  (x, y)
// of a lane's id is advanced by 64 ids per round instead of divided out.
template <int SH, class T>
__device__ __forceinline__ uint32_t compact(const uint32_t *bits, int npix, int dimx, T *list, int lane) {
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t ux = (uint32_t)dimx;
    uint32_t y = (uint32_t)lane / ux, x = (uint32_t)lane - y * ux;
    const uint32_t sy = 64u / ux, sx = 64u - sy * ux;  // sx < dimx
    uint32_t total = 0;
    for (int id0 = 0; id0 < npix; id0 += 64) {
        const int id = id0 + lane;
        const bool in = id < npix && ((bits[id >> 5] >> (id & 31)) & 1u);
        const uint64_t m = __ballot(in);
        if (in) list[total + (uint32_t)__popcll(m & below)] = pack_xy<SH, T>(x, y);
        total += (uint32_t)__popcll(m);
        x += sx;
        y += sy;
        if (x >= ux) {
            x -= ux;
            y++;
        }
    }
    wave_sync();
    return uni(total);
}

__device__ __forceinline__ void clear_bits(uint32_t *bits, int npix, int lane) {
    const int nw = (npix + 31) >> 5;
    for (int w = lane; w < nw; w += 64) bits[w] = 0u;
    wave_sync();
}

__device__ __forceinline__ void set_bit(uint32_t *bits, uint32_t id) { atomicOr(&bits[id >> 5], 1u << (id & 31)); }

struct PairOut {
    uint8_t bit;      // 0 / 1; 3 = the pair does not fit this path's buffers
    double min_dist;
    uint32_t ncand;
    uint64_t evals;
    uint64_t boxes;   // box-pair distances evaluated in the descent
};

// Descend one side: children (at layer g) of every parent in list (at layer
// pg), or (same layer) the members themselves; BORDER ones go to bits.
template <int SH, class T>
__device__ __forceinline__ void expand_side(const T *list, uint32_t n, bool next, const LGrid &pg,
                                            const LGrid &g, const uint8_t *st, uint32_t *bits, int lane) {
    // get_offset_x/_y's (v - low) / step through the step's reciprocal
    // (div_rcp: the same rounded quotient)
    const double ystep = next ? 1.0 / g.step : 0.0;
    auto off = [&](double v, double lo, int dim) {
        return min(max(d2i(div_rcp(v - lo, g.step, ystep)), 0), dim - 1);
    };
    for (uint32_t k = lane; k < n; k += 64) {
        const T v = list[k];
        const int px = unpack_x<SH>(v), py = unpack_y<SH>(v);
        if (!next) {
            const uint32_t id = (uint32_t)(py * g.dimx + px);
            if (border(st[id])) set_bit(bits, id);
            continue;
        }
        // the parent box shrunk by 1e-6, then retrieve_pixels'
        // own 1e-6 (MyRaster.cpp:332-346), in that order
        double b[4];
        pbox(pg, px, py, b);
        b[0] += 1e-6;
        b[1] += 1e-6;
        b[2] -= 1e-6;
        b[3] -= 1e-6;
        const int sx = off(b[0] + 1e-6, g.lx, g.dimx);
        const int sy = off(b[1] + 1e-6, g.ly, g.dimy);
        const int ex = off(b[2] - 1e-6, g.lx, g.dimx);
        const int ey = off(b[3] - 1e-6, g.ly, g.dimy);
        for (int x = sx; x <= ex; x++)
            for (int y = sy; y <= ey; y++) {
                const uint32_t id = (uint32_t)(y * g.dimx + x);
                if (border(st[id])) set_bit(bits, id);
            }
    }
    wave_sync();
}

// Edge-sequence pairs of one candidate pixel pair, 64 at a time.
struct SeqTab {
    uint32_t fa[64], la[64], fb[64], lb[64], w1[64], pre[65];
    unsigned long long mins[64];  // squared distance bits (non-negative doubles order as integers)
    KmC km;                       // the km table entries of the current pair
    Cand stage[64];               // last-level candidates waiting to be merged into the best 64
};

constexpr int kBatch = 4;  // candidates refined together (their sequence pairs share one table)

// A run of candidates taken in tuple order: pixel-pair sequence ranges and
// each one's reference_dist.  All values are uniform across the wave.
struct CandBatch {
    int n = 0;
    uint32_t nrows = 0;
    uint32_t s0[kBatch], t0[kBatch], n1[kBatch], n2[kBatch], row0[kBatch];
    double ref[kBatch];
};

// floor(n / m) for n < 2^22 and 1 <= m: the binary32 quotient is within one
// of it (relative error below 2^-22), and one step against the remainder
// corrects it.
__device__ __forceinline__ uint32_t udiv_small(uint32_t n, uint32_t m) {
    uint32_t q = (uint32_t)((float)n * __builtin_amdgcn_rcpf((float)m));
    const int r = (int)n - (int)(q * m);
    if (r < 0) q--;
    else if (r >= (int)m) q++;
    return q;
}

// Refines a batch of candidate pixel pairs: the truncated distances of all
// their edge-sequence pairs (segment_to_segment_within_batch with vs1 = the
// target sequence, vs2 = the source one), rows = the candidates' sequence
// pairs in order, 64 rows at a time, their evaluations flattened over the
// lanes and each row's minimum gathered in LDS; then the reference's checks
// after every valid sequence pair (running minimum < d: within; < the row's
// candidate's reference_dist: not within) become a running-minimum scan and
// the first lane that passes one.  Returns 1 / 0 when decided, 2 to go on.
__device__ __forceinline__ int refine_batch(const DevSet &S, const farm_raster_hdr &hs, const farm_raster_hdr &ht,
                                            const CandBatch &cb, double d, double &min_dist, PairOut &r,
                                            SeqTab &tab, int lane) {
    const uint32_t nrows = cb.nrows;
    const float *sxv = S.x + hs.vert_start, *syv = S.y + hs.vert_start;
    const float *txv = S.x + ht.vert_start, *tyv = S.y + ht.vert_start;
    for (uint32_t c0 = 0; c0 < nrows; c0 += 64) {
        const uint32_t k = c0 + lane;
        uint32_t la = 0, lb = 0, fa = 0, fb = 0;
        double ref = kFar;
        if (k < nrows) {
            uint32_t s0 = 0, t0 = 0, n2 = 1, rr = 0;
#pragma unroll
            for (int q = 0; q < kBatch; q++) {
                if (q < cb.n && k >= cb.row0[q]) {
                    s0 = cb.s0[q];
                    t0 = cb.t0[q];
                    n2 = cb.n2[q];
                    rr = cb.row0[q];
                    ref = cb.ref[q];
                }
            }
            const uint32_t loc = k - rr, a = loc / n2, b = loc - (loc / n2) * n2;
            const uint2 e1 = S.seqs[hs.seq_start + s0 + a];
            const uint2 e2 = S.seqs[ht.seq_start + t0 + b];
            fa = e1.x;
            la = e1.y;
            fb = e2.x;
            lb = e2.y;
        }
        // sequences of fewer than 2 edges are skipped (Ideal.cpp:2630)
        const bool ok = k < nrows && la >= 2 && lb >= 2;
        const uint32_t w1 = ok ? lb * (la - 1) : 0u;
        const uint32_t w = ok ? w1 + la * (lb - 1) : 0u;
        uint32_t W;
        const uint32_t pre = wave_excl_scan(w, lane, &W);
        tab.fa[lane] = fa;
        tab.la[lane] = la;
        tab.fb[lane] = fb;
        tab.lb[lane] = lb;
        tab.w1[lane] = w1;
        tab.pre[lane] = pre;
        if (lane == 63) tab.pre[64] = W;
        tab.mins[lane] = ~0ull;
        wave_sync();
        // e < w1: point i of the target sequence, segment j of the source
        // one; else point i of the source sequence, segment j of the target
        // one.  One distance per lane either way: the operands are selected,
        // so lanes on both sides of w1 share one evaluation.
        auto sweep = [&](auto small_tag) {
            constexpr bool SM = decltype(small_tag)::value;
            int kk = 0, cur = -1;
            double acc = INFINITY;
            for (uint32_t u = lane; u < W; u += 64) {
                while (tab.pre[kk + 1] <= u) kk++;
                if (kk != cur) {
                    if (cur >= 0) atomicMin(&tab.mins[cur], (unsigned long long)__double_as_longlong(acc));
                    cur = kk;
                    acc = INFINITY;
                }
                const uint32_t e = u - tab.pre[kk];
                const uint32_t fa_ = tab.fa[kk], fb_ = tab.fb[kk], w1_ = tab.w1[kk];
                const bool tgt_pt = e < w1_;
                const uint32_t e2 = tgt_pt ? e : e - w1_, m = (tgt_pt ? tab.la[kk] : tab.lb[kk]) - 1;
                const uint32_t i = SM ? udiv_small(e2, m) : e2 / m, j = e2 - i * m;
                const float *ptx = tgt_pt ? txv + fb_ : sxv + fa_, *pty = tgt_pt ? tyv + fb_ : syv + fa_;
                const float *sgx = tgt_pt ? sxv + fa_ : txv + fb_, *sgy = tgt_pt ? syv + fa_ : tyv + fb_;
                acc = fmin(acc, pt_seg_d2(ptx[i], pty[i], sgx[j], sgy[j], sgx[j + 1], sgy[j + 1], tab.km));
            }
            if (cur >= 0) atomicMin(&tab.mins[cur], (unsigned long long)__double_as_longlong(acc));
        };
        if (W < (1u << 22)) sweep(std::true_type());  // every e2 below 2^22
        else sweep(std::false_type());
        wave_sync();
        // min of the square roots = square root of the min (sqrt is monotone)
        const double dist = ok ? sqrt(__longlong_as_double((long long)tab.mins[lane])) : INFINITY;
        double run = dist;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double t = __shfl_up(run, o, 64);
            if (lane >= o) run = fmin(run, t);
        }
        run = fmin(run, min_dist);
        const unsigned long long hit = __ballot(ok && (run < d || run < ref));
        if (hit) {
            const int ks = uni(__ffsll(hit) - 1);
            const double md = uni(__shfl(run, ks, 64));
            r.evals += tab.pre[ks + 1];
            min_dist = md;
            r.min_dist = md;
            return md < d ? 1 : 0;
        }
        r.evals += W;
        if (__ballot(ok)) {
            min_dist = uni(__shfl(run, 63, 64));
            r.min_dist = min_dist;
        }
        wave_sync();
    }
    return 2;
}

// Argmin (tuple order) over the live candidates, and reference_dist: the
// smallest low bound among the other live candidates (each lane keeps its
// two smallest low bounds).  Lanes feed their candidates (global index k) to
// add(); finish() reduces over the wave and returns the index of the
// smallest, or -1 when none is live.
struct SelAcc {
    Cand best;
    int bi = -1;
    double lo1 = INFINITY, lo2 = INFINITY;
    int i1 = -1;
    __device__ __forceinline__ void add(const Cand &c, int k) {
        if (bi < 0 || tuple_less(c, best)) {
            best = c;
            bi = k;
        }
        if (c.lo < lo1) {
            lo2 = lo1;
            lo1 = c.lo;
            i1 = k;
        } else if (c.lo < lo2) {
            lo2 = c.lo;
        }
    }
    __device__ __forceinline__ int finish(Cand &out, double &ref) {
        if (bi < 0) {
            best.apx = INFINITY;
            best.lo = INFINITY;
            best.hi = INFINITY;
            best.id1 = best.id2 = 0xffffffffu;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            Cand c;
            c.apx = __shfl_xor(best.apx, o, 64);
            c.lo = __shfl_xor(best.lo, o, 64);
            c.hi = __shfl_xor(best.hi, o, 64);
            c.id1 = __shfl_xor(best.id1, o, 64);
            c.id2 = __shfl_xor(best.id2, o, 64);
            const int ci = __shfl_xor(bi, o, 64);
            if (ci >= 0 && (bi < 0 || tuple_less(c, best))) {
                best = c;
                bi = ci;
            }
        }
        out.apx = uni(best.apx);
        out.lo = uni(best.lo);
        out.hi = uni(best.hi);
        out.id1 = uni(best.id1);
        out.id2 = uni(best.id2);
        const int b = uni(bi);
        ref = fmin(kFar, wave_min(i1 == b ? lo2 : lo1));
        return b;
    }
};

// Selection over a candidate slab; a taken candidate has apx = +inf.
__device__ __forceinline__ int select_next(const Cand *cand, uint32_t n, int lane, Cand &out, double &ref) {
    SelAcc acc;
    for (uint32_t k = lane; k < n; k += 64) {
        const Cand c = cand[k];
        if (c.apx == INFINITY) continue;  // taken
        acc.add(c, (int)k);
    }
    return acc.finish(out, ref);
}




__device__ __forceinline__ uint32_t bperm_u(int addr, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)v);
}
__device__ __forceinline__ double bperm_d(int addr, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = bperm_u(addr, (uint32_t)b), hi = bperm_u(addr, (uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One compare-exchange step of a bitonic network over the wave: element i
// (held by this lane) meets element i ^ j (lane ^ j); the lower one keeps
// the smaller tuple when the run is ascending ((i & kk) == 0), the larger
// otherwise.  Equal tuples (only the +inf padding) stay where they are.
__device__ __forceinline__ void bitonic_step(Cand &x, uint32_t i, uint32_t kk, int j) {
    Cand p;
    p.apx = __shfl_xor(x.apx, j, 64);
    p.lo = __shfl_xor(x.lo, j, 64);
    p.hi = __shfl_xor(x.hi, j, 64);
    p.id1 = __shfl_xor(x.id1, j, 64);
    p.id2 = __shfl_xor(x.id2, j, 64);
    const bool up = (i & kk) == 0, low = (i & (uint32_t)j) == 0;
    if (low == up ? tuple_less(p, x) : tuple_less(x, p)) x = p;
}

// The in-register part of a bitonic network for the 64 elements b .. b + 63
// (lane l holds b + l): stages kk = k0 .. 64 (k0 = 2: the whole 64-element
// sort) with their strides below 64, or, for k0 > 64, the strides 32 .. 1 of
// stage k0.  Loops, not unrolled: the kernel's code stays small enough for
// the instruction cache.
__device__ __forceinline__ void bitonic_block(Cand &x, uint32_t b, uint32_t k0, int lane) {
    const uint32_t i = b + (uint32_t)lane;
    if (k0 > 64) {
#pragma unroll 1
        for (int j = 32; j > 0; j >>= 1) bitonic_step(x, i, k0, j);
        return;
    }
#pragma unroll 1
    for (uint32_t kk = k0; kk <= 64; kk <<= 1)
#pragma unroll 1
        for (int j = (int)kk >> 1; j > 0; j >>= 1) bitonic_step(x, i, kk, j);
}

__device__ __forceinline__ Cand cand_inf() {
    Cand c;
    c.apx = c.lo = c.hi = INFINITY;
    c.id1 = c.id2 = 0xffffffffu;
    return c;
}

// Ideal::within for one pair; the wave's buffers: lists / bits (LDS or a
// global slab) and, on the global path, the candidate slab.
template <int SH, class T>
__device__ __forceinline__ PairOut within_pair(const DevSet &S, uint32_t ia, uint32_t ib, double d, T *sl,
                                               T *tl, uint32_t *sb, uint32_t *tb, Cand *cand,
                                               uint32_t pixcap, uint32_t candcap, SeqTab &tab, int lane) {
    PairOut r;
    const double d_le2 = sq_le(d);  // min distance <= d <=> its square <= d_le2
    r.bit = 0;
    r.min_dist = kFar;
    r.ncand = 0;
    r.evals = 0;
    r.boxes = 0;
    const farm_raster_hdr &hs = S.hdr[ia];
    const farm_raster_hdr &ht = S.hdr[ib];
    const int s_level = hs.nlayers - 1, t_level = ht.nlayers - 1;
    {
        // the km table entries of the pair's latitudes (its two grid boxes)
        const double ylo = fmin(hs.mbr[1], ht.mbr[1]), yhi = fmax(hs.mbr[3], ht.mbr[3]);
        const int i0 = (ylo <= 0 && yhi >= 0) ? 0 : km_idx(fmin(fabs(ylo), fabs(yhi)));
        const int i1 = km_idx(fmax(fabs(ylo), fabs(yhi)));
        const int n = min(i1 - i0 + 1, kKmC);
        if (lane < n) {
            tab.km.v[lane] = c_km_lon[i0 + lane];
            tab.km.inv[lane] = c_km_lon_inv[i0 + lane];
        }
        if (lane == 0) {
            tab.km.lo = i0;
            tab.km.n = n;
        }
        wave_sync();
    }
    // every layer must fit the pixel buffers (the finest is the largest,
    // but check all: a coarse layer may be one pixel wider)
    for (int k = 0; k <= s_level; k++) {
        const LGrid g = lgrid(S, hs, k);
        if ((uint64_t)g.dimx * g.dimy > pixcap || g.dimx > (1 << SH) || g.dimy > (1 << SH)) {
            r.bit = 3;
            return r;
        }
    }
    for (int k = 0; k <= t_level; k++) {
        const LGrid g = lgrid(S, ht, k);
        if ((uint64_t)g.dimx * g.dimy > pixcap || g.dimx > (1 << SH) || g.dimy > (1 << SH)) {
            r.bit = 3;
            return r;
        }
    }
    // level 0: every pixel of both top layers
    uint32_t ns, nt;
    {
        const LGrid gs = lgrid(S, hs, 0), gt = lgrid(S, ht, 0);
        ns = (uint32_t)(gs.dimx * gs.dimy);
        nt = (uint32_t)(gt.dimx * gt.dimy);
        for (uint32_t k = lane; k < ns; k += 64) sl[k] = pack_xy<SH, T>(k % gs.dimx, k / gs.dimx);
        for (uint32_t k = lane; k < nt; k += 64) tl[k] = pack_xy<SH, T>(k % gt.dimx, k / gt.dimx);
        wave_sync();
    }
    int i = 0, j = 0;
    bool final_level = false;
    // the LDS path keeps the last level's grids, lists and bound for its own
    // candidate generation (below); the global path writes its candidates to
    // the slab
    LGrid fgs = lgrid(S, hs, 0), fgt = lgrid(S, ht, 0);
    uint32_t fns = 0, fnt = 0;
    double fkeep2 = -1.0;
    for (;;) {
        const double s_step = lgrid(S, hs, i).step, t_step = lgrid(S, ht, j).step;
        bool s_next = false, t_next = false;
        if (i < s_level && (s_step >= t_step || j >= t_level)) {
            i++;
            s_next = true;
        }
        if (j < t_level && (s_step <= t_step || i >= s_level)) {
            j++;
            t_next = true;
        }
        if (ns == 0 || nt == 0) break;  // the queue of pixel pairs is empty
        final_level = !s_next && !t_next;
        const LGrid gs = lgrid(S, hs, i), gt = lgrid(S, ht, j);
        const int npix_s = gs.dimx * gs.dimy, npix_t = gt.dimx * gt.dimy;
        clear_bits(sb, npix_s, lane);
        clear_bits(tb, npix_t, lane);
        const uint8_t *sts = gs.st, *stt = gt.st;
        expand_side<SH>(sl, ns, s_next, s_next ? lgrid(S, hs, i - 1) : gs, gs, sts, sb, lane);
        expand_side<SH>(tl, nt, t_next, t_next ? lgrid(S, ht, j - 1) : gt, gt, stt, tb, lane);
        ns = compact<SH>(sb, npix_s, gs.dimx, sl, lane);
        nt = compact<SH>(tb, npix_t, gt.dimx, tl, lane);
        if (ns == 0 || nt == 0) break;  // no BORDER pixel on a side: nothing to refine
        // pass 1: any BORDER box pair within d by its farthest corners
        // decides; else the smallest such maximum bounds the useful pairs
        // (on squared distances: the smallest maximum is the root of the
        // smallest square, and it is <= d exactly when some maximum is)
        double mx2 = INFINITY;
        const uint64_t total = (uint64_t)ns * nt;
        r.boxes += total;
        {
            uint32_t a = lane / nt, b = lane - (lane / nt) * nt;
            const uint32_t q = 64 / nt, rr = 64 - q * nt;
            while (a < ns) {
                const T va = sl[a], vb = tl[b];
                double b1[4], b2[4];
                pbox(gs, unpack_x<SH>(va), unpack_y<SH>(va), b1);
                pbox(gt, unpack_x<SH>(vb), unpack_y<SH>(vb), b2);
                mx2 = fmin(mx2, box_max_d2(b1, b2, tab.km));
                a += q;
                b += rr;
                if (b >= nt) {
                    b -= nt;
                    a++;
                }
            }
        }
        const double mxmin = sqrt(wave_min(mx2));
        if (mxmin <= d) {
            r.bit = 1;
            return r;
        }
        const double maxd = fmin(kFar, mxmin);
        // min distance <= d and < maxd, as bounds on the squared distance
        const double keep2 = fmin(d_le2, sq_lt(maxd));
        r.boxes += total;
        if constexpr (SH == 8) {
            if (final_level) {
                // the last level's candidates are generated below, the 64
                // smallest at a time (no slab)
                fgs = gs;
                fgt = gt;
                fns = ns;
                fnt = nt;
                fkeep2 = keep2;
                break;
            }
        }
        // pass 2: keep pairs with min distance <= d and below maxd; their
        // projections are the next level's sets, or (last level) they are
        // the candidates
        clear_bits(sb, npix_s, lane);
        clear_bits(tb, npix_t, lane);
        uint32_t nc = 0;
        bool over = false;
        const double sa = gs.step * gs.step, ta = gt.step * gt.step;
        {
            uint32_t a = lane / nt, b = lane - (lane / nt) * nt;
            const uint32_t q = 64 / nt, rr = 64 - q * nt;
            const uint64_t rounds = (total + 63) / 64;
            for (uint64_t it = 0; it < rounds; it++) {
                bool keep = false;
                Cand c;
                c.apx = c.lo = c.hi = 0.0;
                c.id1 = c.id2 = 0;
                if (a < ns) {
                    const T va = sl[a], vb = tl[b];
                    const int xa = unpack_x<SH>(va), ya = unpack_y<SH>(va);
                    const int xb = unpack_x<SH>(vb), yb = unpack_y<SH>(vb);
                    double b1[4], b2[4];
                    pbox(gs, xa, ya, b1);
                    pbox(gt, xb, yb, b2);
                    const double mn2 = box_min_d2(b1, b2, tab.km);
                    if (mn2 <= keep2) {
                        keep = true;
                        const uint32_t id1 = (uint32_t)(ya * gs.dimx + xa), id2 = (uint32_t)(yb * gt.dimx + xb);
                        if (SH == 8 || !final_level) {
                            set_bit(sb, id1);
                            set_bit(tb, id2);
                        } else {
                            const double mn = sqrt(mn2);
                            const double hi = sqrt(box_max_d2(b1, b2, tab.km));
                            const int pf = classify(S.areas[hs.pix_start + id1] + S.areas[ht.pix_start + id2],
                                                    sa + ta, 20);
                            c.apx = mn + (double)c_mean[pf] * (hi - mn);
                            c.lo = mn;
                            c.hi = hi;
                            c.id1 = id1;
                            c.id2 = id2;
                        }
                    }
                }
                if (SH != 8 && final_level) {
                    // global path: every candidate to the slab, in order
                    const uint64_t m = __ballot(keep);
                    const uint32_t pos = nc + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    if (keep) {
                        if (pos < candcap) cand[pos] = c;
                        else over = true;
                    }
                    nc += (uint32_t)__popcll(m);
                }
                a += q;
                b += rr;
                if (b >= nt) {
                    b -= nt;
                    a++;
                }
            }
        }
        wave_sync();
        if (final_level) {
            if (wave_any(over)) {
                r.bit = 3;
                return r;
            }
            r.ncand = nc;
            break;
        }
        ns = compact<SH>(sb, npix_s, gs.dimx, sl, lane);
        nt = compact<SH>(tb, npix_t, gt.dimx, tl, lane);
    }
    if (!final_level) return r;
    if constexpr (SH == 8) {
        // Last level of the LDS path.  A generation pass runs over the final
        // lists' pixel pairs, keeps the candidates (min distance <= d and
        // below the smallest maximum) whose tuple follows `after`, and merges
        // them 64 at a time through the LDS stage into the 64 smallest
        // (sorted, candidate k on lane k); the others only leave their low
        // bound (min_rest).  The refinement takes those 64 in tuple order;
        // when all are refined without a decision and more candidates remain,
        // the next pass starts after the last of them.  Every candidate's
        // reference_dist (the smallest low bound of the candidates after it,
        // Ideal.cpp:2611-2615) is the suffix minimum over the 64 and min_rest.
        const double sa = fgs.step * fgs.step, ta = fgt.step * fgt.step;
        Cand after;
        after.apx = after.lo = after.hi = -INFINITY;
        after.id1 = after.id2 = 0u;
        double min_dist = kFar;
        bool first_pass = true;
        for (;;) {
            Cand best = cand_inf();
            uint32_t nb = 0, count = 0, nst = 0;
            bool full = false;
            Cand th = cand_inf();     // the 64th smallest so far (uniform), once full
            double mrest = INFINITY;  // this lane's smallest low bound outside the best
            auto merge = [&](uint32_t n) {
                wave_sync();
                Cand x = (uint32_t)lane < n ? tab.stage[lane] : cand_inf();
                bitonic_block(x, 0, 2, lane);
                if (nb == 0) {
                    best = x;
                } else {
                    // best ascending, x reversed (descending): the elementwise
                    // smaller is a bitonic sequence of the 64 smallest
                    Cand y;
                    const int src = (63 - lane) << 2;
                    y.apx = bperm_d(src, x.apx);
                    y.lo = bperm_d(src, x.lo);
                    y.hi = bperm_d(src, x.hi);
                    y.id1 = bperm_u(src, x.id1);
                    y.id2 = bperm_u(src, x.id2);
                    const bool ys = tuple_less(y, best);
                    const Cand out = ys ? best : y;  // the larger leaves the 64
                    if (out.apx != INFINITY) mrest = fmin(mrest, out.lo);
                    if (ys) best = y;
#pragma unroll 1
                    for (int j = 32; j > 0; j >>= 1) bitonic_step(best, (uint32_t)lane, 64u, j);
                }
                nb = min(64u, nb + n);
                if (nb == 64) {
                    full = true;
                    th.apx = uni(__shfl(best.apx, 63, 64));
                    th.lo = uni(__shfl(best.lo, 63, 64));
                    th.hi = uni(__shfl(best.hi, 63, 64));
                    th.id1 = uni(__shfl(best.id1, 63, 64));
                    th.id2 = uni(__shfl(best.id2, 63, 64));
                }
                wave_sync();
            };
            const uint64_t total = (uint64_t)fns * fnt;
            if (!first_pass) r.boxes += total;
            {
                uint32_t a = lane / fnt, b = lane - (lane / fnt) * fnt;
                const uint32_t q = 64 / fnt, rr = 64 - q * fnt;
                const uint64_t rounds = (total + 63) / 64;
                for (uint64_t it = 0; it < rounds; it++) {
                    bool keep = false;
                    Cand c = cand_inf();
                    if (a < fns) {
                        const T va = sl[a], vb = tl[b];
                        const int xa = unpack_x<SH>(va), ya = unpack_y<SH>(va);
                        const int xb = unpack_x<SH>(vb), yb = unpack_y<SH>(vb);
                        double b1[4], b2[4];
                        pbox(fgs, xa, ya, b1);
                        pbox(fgt, xb, yb, b2);
                        const double mn2 = box_min_d2(b1, b2, tab.km);
                        if (mn2 <= fkeep2) {
                            const uint32_t id1 = (uint32_t)(ya * fgs.dimx + xa), id2 = (uint32_t)(yb * fgt.dimx + xb);
                            const double mn = sqrt(mn2);
                            const double hi = sqrt(box_max_d2(b1, b2, tab.km));
                            const int pf = classify(S.areas[hs.pix_start + id1] + S.areas[ht.pix_start + id2],
                                                    sa + ta, 20);
                            c.apx = mn + (double)c_mean[pf] * (hi - mn);
                            c.lo = mn;
                            c.hi = hi;
                            c.id1 = id1;
                            c.id2 = id2;
                            keep = tuple_less(after, c);
                        }
                    }
                    count += (uint32_t)__popcll(__ballot(keep));
                    const bool enter = keep && (!full || tuple_less(c, th));
                    if (keep && !enter) mrest = fmin(mrest, c.lo);
                    const uint64_t m = __ballot(enter);
                    const uint32_t cnt = (uint32_t)__popcll(m);
                    if (nst + cnt > 64) {
                        merge(nst);
                        nst = 0;
                    }
                    if (enter) tab.stage[nst + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = c;
                    nst += cnt;
                    a += q;
                    b += rr;
                    if (b >= fnt) {
                        b -= fnt;
                        a++;
                    }
                }
            }
            if (nst) merge(nst);
            if (first_pass) {
                r.ncand = count;
                first_pass = false;
            }
            if (count == 0) return r;
            // reference_dist of lane k's candidate, and its pixel ids
            const double rest = wave_min(mrest);
            double sfx = (uint32_t)lane < nb ? best.lo : INFINITY;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double t = __shfl_down(sfx, o, 64);
                if (lane + o < 64) sfx = fmin(sfx, t);
            }
            const double later = __shfl_down(sfx, 1, 64);
            const double rref = fmin(kFar, fmin(lane < 63 ? later : INFINITY, rest));
            const uint32_t rids = (best.id1 & 0xffffu) | (best.id2 << 16);
            uint32_t k = 0;
            while (k < nb) {
                CandBatch cb;
                bool open = true;
#pragma unroll
                for (int qq = 0; qq < kBatch; qq++) {
                    if (open && k < nb) {
                        const int kk = (int)k;
                        const uint32_t ids = (uint32_t)__builtin_amdgcn_readlane((int)rids, kk);
                        const uint32_t id1 = ids & 0xffffu, id2 = ids >> 16;
                        const long long rb = __double_as_longlong(rref);
                        const int rlo = __builtin_amdgcn_readlane((int)(rb & 0xffffffffll), kk);
                        const int rhi = __builtin_amdgcn_readlane((int)(rb >> 32), kk);
                        const double ref = __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
                        const uint32_t s0 = S.pixoff[hs.pixoff_start + id1], s1 = S.pixoff[hs.pixoff_start + id1 + 1];
                        const uint32_t t0 = S.pixoff[ht.pixoff_start + id2], t1 = S.pixoff[ht.pixoff_start + id2 + 1];
                        const uint32_t rows = (s1 - s0) * (t1 - t0);
                        if (qq > 0 && cb.nrows + rows > 64) {
                            open = false;  // candidate k starts the next run
                        } else {
                            cb.s0[qq] = s0;
                            cb.t0[qq] = t0;
                            cb.n1[qq] = s1 - s0;
                            cb.n2[qq] = t1 - t0;
                            cb.ref[qq] = ref;
                            cb.row0[qq] = cb.nrows;
                            cb.nrows += rows;
                            cb.n = qq + 1;
                            k++;
                            if (cb.nrows >= 64) open = false;
                        }
                    }
                }
                const int res = refine_batch(S, hs, ht, cb, d, min_dist, r, tab, lane);
                if (res != 2) {
                    r.bit = (uint8_t)res;
                    return r;
                }
            }
            if (count <= nb) return r;  // every candidate refined: not within
            after.apx = uni(__shfl(best.apx, 63, 64));
            after.lo = uni(__shfl(best.lo, 63, 64));
            after.hi = uni(__shfl(best.hi, 63, 64));
            after.id1 = uni(__shfl(best.id1, 63, 64));
            after.id2 = uni(__shfl(best.id2, 63, 64));
        }
    }
    if constexpr (SH != 8) {
        if (r.ncand == 0) return r;
        // global path: the slab's candidates are taken one at a time by
        // argmin selection in tuple order (with each one's reference_dist)
        // and refined in runs of up to kBatch whose sequence pairs fit one
        // 64-row table (a candidate with more rows runs alone, 64 rows at a
        // time)
        const uint32_t nc = r.ncand;
        double min_dist = kFar;
        Cand c;
        double ref;
        int slot = select_next(cand, nc, lane, c, ref);
        while (slot >= 0) {
            CandBatch cb;
            bool open = true;
#pragma unroll
            for (int q = 0; q < kBatch; q++) {
                if (open && slot >= 0) {
                    const uint32_t s0 = S.pixoff[hs.pixoff_start + c.id1], s1 = S.pixoff[hs.pixoff_start + c.id1 + 1];
                    const uint32_t t0 = S.pixoff[ht.pixoff_start + c.id2], t1 = S.pixoff[ht.pixoff_start + c.id2 + 1];
                    const uint32_t rows = (s1 - s0) * (t1 - t0);
                    if (q > 0 && cb.nrows + rows > 64) {
                        open = false;  // c starts the next run
                    } else {
                        cb.s0[q] = s0;
                        cb.t0[q] = t0;
                        cb.n1[q] = s1 - s0;
                        cb.n2[q] = t1 - t0;
                        cb.ref[q] = ref;
                        cb.row0[q] = cb.nrows;
                        cb.nrows += rows;
                        cb.n = q + 1;
                        if (lane == 0) cand[slot].apx = INFINITY;
                        wave_sync();
                        slot = select_next(cand, nc, lane, c, ref);
                        if (cb.nrows >= 64) open = false;
                    }
                }
            }
            const int res = refine_batch(S, hs, ht, cb, d, min_dist, r, tab, lane);
            if (res != 2) {
                r.bit = (uint8_t)res;
                return r;
            }
        }
    }
    return r;
}

}  // namespace

// Persistent waves take pairs in chunks from a global counter.  BIG: the
// pair indices come from `list` (pairs the LDS path could not hold) and the
// lists / bits live in a per-wave global slab.
template <bool BIG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_within(DevSet S, const uint32_t *__restrict__ pa,
                                                const uint32_t *__restrict__ pb, uint32_t n,
                                                const uint32_t *__restrict__ list, double d, uint8_t *__restrict__ out,
                                                double *__restrict__ md, char *__restrict__ slab, uint64_t slab_stride,
                                                uint32_t pixcap, uint32_t candcap, uint32_t *__restrict__ ctr,
                                                uint32_t *__restrict__ ovf, unsigned long long *__restrict__ stats) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const uint64_t gw = (uint64_t)blockIdx.x * kWaves + wv;
    __shared__ SeqTab tabs[kWaves];
    using LT = typename std::conditional<BIG, uint32_t, uint16_t>::type;
    constexpr int SH = BIG ? 16 : 8;
    LT *sl, *tl;
    uint32_t *sb, *tb;
    Cand *cand;
    if constexpr (!BIG) {
        __shared__ LT lists[kWaves][2][kPixMax];
        __shared__ uint32_t bits[kWaves][2][kBitWords];
        sl = lists[wv][0];
        tl = lists[wv][1];
        sb = bits[wv][0];
        tb = bits[wv][1];
        cand = nullptr;  // the LDS path keeps its candidates in registers and LDS
    } else {
        char *base = slab + gw * slab_stride;
        const uint64_t wbits = ((uint64_t)pixcap + 31) / 32;
        sl = reinterpret_cast<LT *>(base);
        tl = sl + pixcap;
        sb = reinterpret_cast<uint32_t *>(tl + pixcap);
        tb = sb + wbits;
        cand = reinterpret_cast<Cand *>(base + (((2ull * pixcap + 2 * wbits) * 4 + 255) & ~255ull));
    }
    unsigned long long ncand = 0, evals = 0, found = 0, boxes = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&ctr[0], kChunk);
        base = uni(__shfl(base, 0, 64));
        if (base >= n) break;
        const uint32_t end = min(base + kChunk, n);
        for (uint32_t k = base; k < end; k++) {
            const uint32_t idx = BIG ? list[k] : k;
            const uint32_t ia = pa[idx], ib = pb[idx];
            const farm_raster_hdr &hs = S.hdr[ia];
            const farm_raster_hdr &ht = S.hdr[ib];
            if (((hs.flags | ht.flags) & FARM_HDR_REJECTED) || hs.nlayers <= 0 || ht.nlayers <= 0) {
                if (lane == 0) {
                    out[idx] = 2;
                    md[idx] = kFar;
                }
                continue;
            }
            const PairOut r = within_pair<SH>(S, ia, ib, d, sl, tl, sb, tb, cand, pixcap, candcap, tabs[wv], lane);
            if (r.bit == 3) {
                if (lane == 0) {
                    if (BIG) {
                        out[idx] = 3;
                    } else {
                        const uint32_t slot = atomicAdd(&ctr[1], 1u);
                        ovf[slot] = idx;
                    }
                }
                continue;
            }
            ncand += r.ncand;
            evals += r.evals;
            boxes += r.boxes;
            found += r.bit;
            if (lane == 0) {
                out[idx] = r.bit;
                md[idx] = r.min_dist;
            }
        }
    }
    if (lane == 0) {
        atomicAdd(&stats[0], ncand);
        atomicAdd(&stats[1], evals);
        atomicAdd(&stats[2], found);
        atomicAdd(&stats[3], boxes);
    }
}

template __global__ void k_within<false>(DevSet, const uint32_t *, const uint32_t *, uint32_t, const uint32_t *, double,
                                         uint8_t *, double *, char *, uint64_t, uint32_t, uint32_t, uint32_t *,
                                         uint32_t *, unsigned long long *);
template __global__ void k_within<true>(DevSet, const uint32_t *, const uint32_t *, uint32_t, const uint32_t *, double,
                                        uint8_t *, double *, char *, uint64_t, uint32_t, uint32_t, uint32_t *,
                                        uint32_t *, unsigned long long *);

}  // namespace farm

using namespace farm;

namespace {

int within_consts(farm_handle h) {
    if (h->within_consts_ready) return FARM_OK;
    static const float mean[20] = {0.0f,  0.7f,  0.62f, 0.55f, 0.52f, 0.45f, 0.42f, 0.4f,  0.38f, 0.34f,
                                   0.31f, 0.28f, 0.25f, 0.22f, 0.18f, 0.15f, 0.1f,  0.05f, 0.0f,  0.0f};
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_km_lon), km_lon_table(), sizeof(double) * 900));
    double inv[900];
    for (int i = 0; i < 900; i++) inv[i] = 1.0 / km_lon_table()[i];
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_km_lon_inv), inv, sizeof inv));
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_mean), mean, sizeof mean));
    h->within_consts_ready = true;
    return FARM_OK;
}

int run_within(farm_handle h, uint64_t begin, uint64_t end, double d_km, uint8_t *out_dev, double *md_dev,
               farm_stats *st) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    if (!h->has_layers) return fail(h, FARM_ESTATE, "within needs a set rasterized with layers (-h)");
    if (begin > end || end > h->npairs)
        return fail(h, FARM_EINVAL, "pair window [%llu,%llu) outside [0,%llu)", (unsigned long long)begin,
                    (unsigned long long)end, (unsigned long long)h->npairs);
    if (!(d_km >= 0)) return fail(h, FARM_EINVAL, "distance must be >= 0");
    const uint64_t n64 = end - begin;
    if (n64 > 0xffffffffull) return fail(h, FARM_ECAPACITY, "batch larger than 2^32 pairs");
    const uint32_t n = (uint32_t)n64;
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = within_consts(h))) return rc;
    if (st) memset(st, 0, sizeof *st);
    if (st) st->pairs = n;
    if (n == 0) return FARM_OK;
    const uint32_t *pa = P<uint32_t>(h->pa) + begin;
    const uint32_t *pb = P<uint32_t>(h->pb) + begin;
    // counters: [0] queue head, [1] overflow count, [2..3] big-path head; stats
    if ((rc = ensure(h, h->wctr, 64))) return rc;
    if ((rc = ensure(h, h->wovf, 4ull * n))) return rc;
    uint32_t *ctr = P<uint32_t>(h->wctr);
    unsigned long long *wst = reinterpret_cast<unsigned long long *>(ctr + 8);
    if ((rc = fill_u32(h, ctr, 16, 0))) return rc;
    // as many resident workgroups as the registers and LDS allow (the waves
    // are persistent)
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_within<false>, 64 * kWaves, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    const unsigned nblocks = (unsigned)h->num_cus * (unsigned)std::min(per_cu, 8);
    DevSet S = devset(h);
    HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    hipLaunchKernelGGL(k_within<false>, dim3(nblocks), dim3(64 * kWaves), 0, h->stream, S, pa, pb, n,
                       (const uint32_t *)nullptr, d_km, out_dev, md_dev, (char *)nullptr, (uint64_t)0,
                       (uint32_t)kPixMax, (uint32_t)0, ctr, P<uint32_t>(h->wovf), wst);
    HIPCHK(h, hipGetLastError());
    uint32_t novf = 0;
    HIPCHK(h, hipMemcpyAsync(&novf, ctr + 1, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (novf) {
        // pairs whose layers or candidates exceed the LDS path: the same
        // algorithm with per-wave global slabs sized for the largest layer
        // of the set and a candidate bound of the border pixel counts
        const uint32_t pixcap = (uint32_t)std::max<uint64_t>(h->max_layer_pix, 1);
        const uint64_t candcap = std::min<uint64_t>(h->max_border_pix * h->max_border_pix + 64, 1ull << 24);
        const uint64_t wbits = (pixcap + 31) / 32;
        const uint64_t head = (((2ull * pixcap + 2 * wbits) * 4 + 255) & ~255ull);
        const uint64_t bstride = head + ((candcap * sizeof(Cand) + 255) & ~255ull);
        uint64_t bwaves = std::min<uint64_t>((uint64_t)h->num_cus * 4, std::max<uint64_t>(1, (4ull << 30) / bstride));
        bwaves = std::min<uint64_t>(bwaves, ((uint64_t)novf + kWaves - 1) / kWaves * kWaves);
        const unsigned bblocks = (unsigned)std::max<uint64_t>(1, bwaves / kWaves);
        if ((rc = ensure(h, h->wslab_big, (uint64_t)bblocks * kWaves * bstride))) return rc;
        if ((rc = fill_u32(h, ctr, 1, 0))) return rc;
        hipLaunchKernelGGL(k_within<true>, dim3(bblocks), dim3(64 * kWaves), 0, h->stream, S, pa, pb, novf,
                           P<uint32_t>(h->wovf), d_km, out_dev, md_dev, P<char>(h->wslab_big), bstride, pixcap,
                           (uint32_t)candcap, ctr, P<uint32_t>(h->wovf), wst);
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    unsigned long long hst[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(hst, wst, sizeof hst, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (novf) {
        // a pair that even the global slabs cannot hold is an error, not a
        // silent answer
        std::vector<uint8_t> tmp(n);
        HIPCHK(h, hipMemcpy(tmp.data(), out_dev, n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; k++)
            if (tmp[k] == 3)
                return fail(h, FARM_ECAPACITY, "pair %llu: more candidate pixel pairs than the within buffers hold",
                            (unsigned long long)(begin + k));
    }
    if (st) {
        float ms = 0;
        HIPCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        st->bb_pixel_pairs = hst[0];
        st->segment_tests = hst[1];
        st->found = hst[2];
        st->refine_items = hst[3];
        st->refine_ms = ms;
        st->total_ms = ms;
    }
    return FARM_OK;
}

}  // namespace

extern "C" {

int farm_within_dev(farm_handle h, uint64_t begin, uint64_t end, double d_km, uint8_t *out_dev, double *md_dev,
                    farm_stats *st) {
    if (!out_dev || !md_dev) return fail(h, FARM_EINVAL, "null output buffer");
    return run_within(h, begin, end, d_km, out_dev, md_dev, st);
}

int farm_within(farm_handle h, uint64_t begin, uint64_t end, double d_km, uint8_t *out, double *min_dist,
                farm_stats *st) {
    if (!h) return FARM_EINVAL;
    if (begin > end) return fail(h, FARM_EINVAL, "bad pair window");
    const uint64_t n = end - begin;
    int rc;
    if ((rc = ensure(h, h->wout, n + 16))) return rc;
    if ((rc = ensure(h, h->wmd, 8 * n + 16))) return rc;
    // out / min_dist may be NULL: results stay on the device (counts only)
    if ((rc = run_within(h, begin, end, d_km, P<uint8_t>(h->wout), P<double>(h->wmd), st))) return rc;
    if (n && out) HIPCHK(h, hipMemcpy(out, h->wout.p, n, hipMemcpyDeviceToHost));
    if (n && min_dist) HIPCHK(h, hipMemcpy(min_dist, h->wmd.p, 8 * n, hipMemcpyDeviceToHost));
    return FARM_OK;
}

}  // extern "C"
