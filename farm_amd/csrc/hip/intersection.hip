// Intersection query on gfx950: boundary crossing points, classified
// boundary pieces and the area of A n B.
//
// Follows Ideal::intersection (src/geometry/Ideal.cpp:1440-1683) up to the
// piece statuses, where the reference returns: the crossing points of edges
// in BORDER x BORDER pixel pairs (segment_intersect_batch,
// src/include/geometry_computation.h:239-267), sorted and de-duplicated by
// point, the boundary of each polygon split at them, and every piece
// classified by Ideal::segment_contain (Ideal.cpp:1233-1327).  On top of
// those pieces the area is integrated with Green's theorem:
//   area = |sA * sum(A pieces IN B) + sB * sum(B pieces IN A)| / 2
// (cross-product sums of the piece paths relative to the low corner of the
// two MBBs' overlap; sX the orientation of ring X; a piece is IN when the
// midpoint of its longest sub-segment is).  Without two distinct crossing
// points the pair is contained (first vertex IN the other) or disjoint.
//
// One wavefront per pair:
//  1. the source raster's window over the target MBB; all OUT / all IN
//     windows end the pair (the reference's early return)
//  2. every edge pair of every BORDER x BORDER pixel pair's sequence pairs:
//     source edges on lanes, target vertices streamed; a line screen (the
//     side of each target vertex against the source edge's line, with a
//     margin above the f32 rounding) drops edge pairs that cannot cross;
//     the rest are queued in LDS and take the reference's exact f32 / f64
//     test a wave at a time; hits are appended to a per-wave record slab in
//     global memory
//  3. bitonic sorts of the slab (by point; by source edge and parameter;
//     by target edge and parameter), de-duplication, pieces, their statuses
//     (point-in-polygon tests, wave-cooperative when a chunk holds few
//     pieces) and the wave-cooperative path sums of the IN pieces
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>

#include "handle.h"
#include "wave.h"

namespace farm {

namespace {

struct Rec {
    float px, py;
    int32_t ei, ej;
    double t, u;
};

constexpr int kWaves = 4;
constexpr uint32_t kRecCap = 2048;     // records per wave slab (power of two)
constexpr uint32_t kRecCapBig = 65536;  // second pass for pairs beyond kRecCap (a third sizes to the largest)
constexpr uint32_t kChunk = 2;          // pairs per queue grab

// Point::operator== (Point.h:32)
__device__ __forceinline__ bool pt_eq(float ax, float ay, float bx, float by) {
    return fabsf(ax - bx) < 1e-9 && fabsf(ay - by) < 1e-9;
}

// key orders: 0 by point (then ei, ej, t, u); 1 by source edge and t (then
// ej, u); 2 by target edge and u (then ei, t)
template <int K>
__device__ __forceinline__ bool rec_less(const Rec &a, const Rec &b) {
    if constexpr (K == 0) {
        if (a.px != b.px) return a.px < b.px;
        if (a.py != b.py) return a.py < b.py;
        if (a.ei != b.ei) return a.ei < b.ei;
        if (a.ej != b.ej) return a.ej < b.ej;
        if (a.t != b.t) return a.t < b.t;
        return a.u < b.u;
    } else if constexpr (K == 1) {
        if (a.ei != b.ei) return a.ei < b.ei;
        if (a.t != b.t) return a.t < b.t;
        if (a.ej != b.ej) return a.ej < b.ej;
        return a.u < b.u;
    } else {
        if (a.ej != b.ej) return a.ej < b.ej;
        if (a.u != b.u) return a.u < b.u;
        if (a.ei != b.ei) return a.ei < b.ei;
        return a.t < b.t;
    }
}

__device__ __forceinline__ Rec pad_rec() {
    Rec r;
    r.px = INFINITY;
    r.py = INFINITY;
    r.ei = INT_MAX;
    r.ej = INT_MAX;
    r.t = INFINITY;
    r.u = INFINITY;
    return r;
}

// Bitonic sort of R[0, n) in place; R has room for next_pow2(n) records.
template <int K>
__device__ void bitonic(Rec *R, uint32_t n, int lane) {
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = n + lane; i < m; i += 64) R[i] = pad_rec();
    wave_sync();
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < m; i += 64) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const Rec a = R[i], b = R[l];
                    const bool up = (i & k) == 0;
                    if (rec_less<K>(b, a) == up) {
                        R[i] = b;
                        R[l] = a;
                    }
                }
            }
            wave_sync();
        }
    }
}

// Ideal::segment_contain with count_intersection_nodes (Ideal.cpp:988-1006):
// 0 OUT, 1 BORDER (on the boundary), 2 IN
__device__ int seg_contain(const DevSet &S, const farm_raster_hdr &h, float px, float py) {
    const int x = off_axis((double)px, h.mbr[0], h.step[0], h.dimx);
    const int y = off_axis((double)py, h.mbr[1], h.step[1], h.dimy);
    const uint32_t id = (uint32_t)(y * h.dimx + x);
    const uint8_t st = S.status[h.pix_start + id];
    if (st == kCC - 1) return 2;
    if (st == 0) return 0;
    const double bx_high = h.mbr[0] + (x + 1) * h.step[0];
    bool ret = false;
    const uint32_t s0 = S.pixoff[h.pixoff_start + id], s1 = S.pixoff[h.pixoff_start + id + 1];
    const float *X = S.x + h.vert_start, *Y = S.y + h.vert_start;
    for (uint32_t e = s0; e < s1; e++) {
        const uint2 sq = S.seqs[h.seq_start + e];
        for (uint32_t k = 0; k < sq.y; k++) {
            const float v1x = X[sq.x + k], v1y = Y[sq.x + k];
            const float v2x = X[sq.x + k + 1], v2y = Y[sq.x + k + 1];
            if (pt_eq(px, py, v1x, v1y) || pt_eq(px, py, v2x, v2y)) return 1;
            if ((v1y >= py) != (v2y >= py)) {
                const double dx = (double)(v2x - v1x);
                const double dy = (double)(v2y - v1y);
                const double pyd = (double)(py - v1y);
                if (fabs(dy) > 1e-9) {
                    const double ix = dx * pyd / dy + (double)v1x;
                    if (fabs((double)px - ix) < 1e-9) return 1;
                    if ((double)px < ix && ix <= bx_high) ret = !ret;
                }
            } else if (v1y == py && v2y == py && ((v1x >= px) != (v2x >= px))) {
                return 1;
            }
        }
    }
    const uint32_t *vo = S.voff + h.vline_start;
    uint32_t i = vo[x + 1];
    const uint32_t j = vo[x + 2];
    int cnt = 0;
    while (i < j && S.vnodes[h.vnode_start + i] < (double)py) {
        cnt++;
        i++;
    }
    if (cnt & 1) ret = !ret;
    return ret ? 2 : 0;
}

// seg_contain for one point (the same for every lane), the whole wave
// testing the pixel's edges together.  The serial form returns 1 at the first
// edge that puts the point on the boundary, else the parity of the crossings
// to the right within the pixel and of the sorted grid-line nodes below it:
// neither depends on the order the edges are visited in, so "on" is a wave
// any, the parity a ballot count, and the node prefix a count of nodes < py.
__device__ int seg_contain_wave(const DevSet &S, const farm_raster_hdr &h, float px, float py, int lane) {
    const int x = off_axis((double)px, h.mbr[0], h.step[0], h.dimx);
    const int y = off_axis((double)py, h.mbr[1], h.step[1], h.dimy);
    const uint32_t id = (uint32_t)(y * h.dimx + x);
    const uint8_t st = S.status[h.pix_start + id];
    if (st == kCC - 1) return 2;
    if (st == 0) return 0;
    const double bx_high = h.mbr[0] + (x + 1) * h.step[0];
    bool on = false, par = false;
    const uint32_t s0 = S.pixoff[h.pixoff_start + id], s1 = S.pixoff[h.pixoff_start + id + 1];
    const float *X = S.x + h.vert_start, *Y = S.y + h.vert_start;
    for (uint32_t e = s0; e < s1; e++) {
        const uint2 sq = S.seqs[h.seq_start + e];
        for (uint32_t k = lane; k < sq.y; k += 64) {
            const float v1x = X[sq.x + k], v1y = Y[sq.x + k];
            const float v2x = X[sq.x + k + 1], v2y = Y[sq.x + k + 1];
            if (pt_eq(px, py, v1x, v1y) || pt_eq(px, py, v2x, v2y)) on = true;
            if ((v1y >= py) != (v2y >= py)) {
                const double dx = (double)(v2x - v1x);
                const double dy = (double)(v2y - v1y);
                const double pyd = (double)(py - v1y);
                if (fabs(dy) > 1e-9) {
                    const double ix = dx * pyd / dy + (double)v1x;
                    if (fabs((double)px - ix) < 1e-9) on = true;
                    if ((double)px < ix && ix <= bx_high) par = !par;
                }
            } else if (v1y == py && v2y == py && ((v1x >= px) != (v2x >= px))) {
                on = true;
            }
        }
    }
    if (wave_any(on)) return 1;
    uint32_t cnt = (uint32_t)__popcll(__ballot(par));
    const uint32_t *vo = S.voff + h.vline_start;
    const uint32_t i = vo[x + 1], j = vo[x + 2];
    for (uint32_t k0 = i; k0 < j; k0 += 64) {
        const uint32_t k = k0 + lane;
        cnt += (uint32_t)__popcll(__ballot(k < j && S.vnodes[h.vnode_start + k] < (double)py));
    }
    return (cnt & 1) ? 2 : 0;
}

// The IN pieces of a pair go to buf (when requested), one atomic per wave
// and chunk; beyond cap they are only counted.
struct PieceOut {
    farm_piece *buf;
    uint32_t *cnt;
    uint32_t cap;
    uint32_t pair;
};

__device__ __forceinline__ void put_pieces(const PieceOut &po, bool mine, const farm_piece &p, int lane) {
    const uint64_t m = __ballot(mine);
    if (!po.buf || !m) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(po.cnt, (uint32_t)__popcll(m));
    base = uni(__shfl(base, 0, 64));
    if (mine) {
        const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (k < po.cap) po.buf[k] = p;
    }
}

struct PairRes {
    int32_t c[5];  // [crossing points, pieces, IN, OUT, BORDER]
    double area;
    uint64_t bb, tests, raw;
    bool over;
};

// Path point k of an area piece: 0 = a, 1..nvert = ring vertices from
// `first` (cyclic over m), nvert + 1 = b.
__device__ __forceinline__ void path_pt(const float *X, const float *Y, int m, int first, int nvert, float ax,
                                        float ay, float bx, float by, int k, float &x, float &y) {
    if (k == 0) {
        x = ax;
        y = ay;
    } else if (k == nvert + 1) {
        x = bx;
        y = by;
    } else {
        int v = first + k - 1;
        if (v >= m) v -= m;
        x = X[v];
        y = Y[v];
    }
}

constexpr int kLanePath = 16;  // area-piece paths of at most this many segments are searched by one lane

// One ring's pieces: D[0, n) sorted by this ring's (edge, parameter).
// Counts the reference pieces by status into r and returns this lane's part
// of the cross-product sum of the area pieces IN the other polygon.
template <bool SRC>
__device__ double ring_pieces(const DevSet &S, const farm_raster_hdr &hp, const farm_raster_hdr &hq, const Rec *D,
                              uint32_t n, double ox, double oy, int lane, PairRes &r, uint32_t &pieces,
                              const PieceOut &po) {
    const float *X = S.x + hp.vert_start, *Y = S.y + hp.vert_start;
    const int m = (int)hp.nverts - 1;
    const double eps = (double)1e-6f;  // util.h:44
    double acc = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        const uint32_t i = c0 + lane;
        bool piece = false, in_area = false, long_path = false;
        int st = 0, first = 0, nvert = 0;
        float ax = 0, ay = 0, bx = 0, by = 0;
        float rx = 0, ry = 0, mx = 0, my = 0;  // the two points tested against the other polygon
        if (i < n) {
            const Rec a = D[i], b = D[(i + 1) % n];
            ax = a.px;
            ay = a.py;
            bx = b.px;
            by = b.py;
            if (!pt_eq(a.px, a.py, b.px, b.py)) {
                piece = true;
                // the reference's segment (Ideal.cpp:1559-1581, 1592-1614)
                const int ea0 = SRC ? a.ei : a.ej, eb0 = SRC ? b.ei : b.ej;
                const double ta0 = SRC ? a.t : a.u, tb0 = SRC ? b.t : b.u;
                int ae = ea0, be = eb0;
                double ap = ta0, bp = tb0;
                if (fabs(ap - 1.0) < eps) {
                    ae = (ae + 1) % m;
                    ap = 0.0;
                }
                if (fabs(bp) < eps) {
                    be--;
                    bp = 1.0;
                }
                if (ae == be && ap < bp) {
                    rx = (a.px + b.px) * 0.5f;
                    ry = (a.py + b.py) * 0.5f;
                } else {
                    rx = X[ae + 1];
                    ry = Y[ae + 1];
                }
                // the area piece: a -> ring vertices ea0 + 1 .. eb0 -> b
                if (i + 1 < n) nvert = eb0 - ea0;
                else nvert = (eb0 - ea0 + m) % m + ((eb0 == ea0 && tb0 <= ta0) ? m : 0);
                if (nvert < 0) nvert = 0;
                first = ea0 + 1;
                // the area's test point: the midpoint of the path's longest
                // sub-segment (the first on ties), well inside the piece; a
                // path of more than kLanePath segments is searched by the
                // whole wave below
                if (nvert + 1 <= kLanePath) {
                    float ux = ax, uy = ay, vx = ax, vy = ay, px = ax, py = ay;
                    double best = -1.0;
                    for (int k = 1; k <= nvert + 1; k++) {
                        float qx, qy;
                        path_pt(X, Y, m, first, nvert, ax, ay, bx, by, k, qx, qy);
                        const double dx = (double)qx - px, dy = (double)qy - py;
                        const double l = dx * dx + dy * dy;
                        if (l > best) {
                            best = l;
                            ux = px;
                            uy = py;
                            vx = qx;
                            vy = qy;
                        }
                        px = qx;
                        py = qy;
                    }
                    mx = (ux + vx) * 0.5f;
                    my = (uy + vy) * 0.5f;
                } else {
                    long_path = true;
                }
            }
        }
        // long paths, one at a time: segment k (points k - 1 -> k) on lane
        // (k - 1) % 64; each lane keeps its first longest, then the wave takes
        // the longest with the smallest k
        uint64_t lp = __ballot(long_path);
        while (lp) {
            const int src = __ffsll((unsigned long long)lp) - 1;
            lp &= lp - 1;
            const int f = uni(__shfl(first, src, 64)), nv = uni(__shfl(nvert, src, 64));
            const float pax = uni(__shfl(ax, src, 64)), pay = uni(__shfl(ay, src, 64));
            const float pbx = uni(__shfl(bx, src, 64)), pby = uni(__shfl(by, src, 64));
            double best = -1.0;
            int bk = INT_MAX;
            float ux = 0.f, uy = 0.f, vx = 0.f, vy = 0.f;
            for (int k = 1 + lane; k <= nv + 1; k += 64) {
                float x0, y0, x1, y1;
                path_pt(X, Y, m, f, nv, pax, pay, pbx, pby, k - 1, x0, y0);
                path_pt(X, Y, m, f, nv, pax, pay, pbx, pby, k, x1, y1);
                const double dx = (double)x1 - x0, dy = (double)y1 - y0;
                const double l = dx * dx + dy * dy;
                if (l > best) {
                    best = l;
                    bk = k;
                    ux = x0;
                    uy = y0;
                    vx = x1;
                    vy = y1;
                }
            }
            const double top = -wave_min(-best);
            int kk = best == top ? bk : INT_MAX;
            for (int o = 32; o > 0; o >>= 1) kk = min(kk, __shfl_xor(kk, o, 64));
            const int wl = __ffsll((unsigned long long)__ballot(best == top && bk == kk)) - 1;
            const float wmx = (__shfl(ux, wl, 64) + __shfl(vx, wl, 64)) * 0.5f;
            const float wmy = (__shfl(uy, wl, 64) + __shfl(vy, wl, 64)) * 0.5f;
            if (lane == src) {
                mx = wmx;
                my = wmy;
            }
        }
        const uint64_t pm = __ballot(piece);
        if (__popcll(pm) <= 32) {
            // few pieces: the wave tests each point's pixel edges together
            uint64_t q = pm;
            while (q) {
                const int src = __ffsll((unsigned long long)q) - 1;
                q &= q - 1;
                const int s1 = seg_contain_wave(S, hq, uni(__shfl(rx, src, 64)), uni(__shfl(ry, src, 64)), lane);
                const int s2 = seg_contain_wave(S, hq, uni(__shfl(mx, src, 64)), uni(__shfl(my, src, 64)), lane);
                if (lane == src) {
                    st = s1;
                    in_area = s2 == 2;
                }
            }
        } else if (piece) {
            st = seg_contain(S, hq, rx, ry);
            in_area = seg_contain(S, hq, mx, my) == 2;
        }
        r.c[1] += __popcll(pm);
        r.c[2] += __popcll(__ballot(piece && st == 2));
        r.c[3] += __popcll(__ballot(piece && st == 0));
        r.c[4] += __popcll(__ballot(piece && st == 1));
        pieces += __popcll(pm);
        put_pieces(po, in_area, farm_piece{po.pair, SRC ? 0 : 1, first, nvert, ax, ay, bx, by}, lane);
        // wave-cooperative path sums of this chunk's IN pieces
        uint64_t inm = __ballot(in_area);
        while (inm) {
            const int src = __ffsll((unsigned long long)inm) - 1;
            inm &= inm - 1;
            const int f = __shfl(first, src, 64), nv = __shfl(nvert, src, 64);
            const float pax = __shfl(ax, src, 64), pay = __shfl(ay, src, 64);
            const float pbx = __shfl(bx, src, 64), pby = __shfl(by, src, 64);
            for (int k = lane; k <= nv; k += 64) {
                float x0, y0, x1, y1;
                path_pt(X, Y, m, f, nv, pax, pay, pbx, pby, k, x0, y0);
                path_pt(X, Y, m, f, nv, pax, pay, pbx, pby, k + 1, x1, y1);
                acc += ((double)x0 - ox) * ((double)y1 - oy) - ((double)x1 - ox) * ((double)y0 - oy);
            }
        }
    }
    return acc;
}

// The reference's crossing test of source edge ea and target edge eb
// (segment_intersect_batch, geometry_computation.h:239-267: f32 edge vectors
// and crosses, f64 parameters, f32 point), behind a screen that drops pairs
// whose parameters c2/dn, c1/dn lie clearly outside [-1e-6, 1 + 1e-6]
// (the f32 products in the screen are off by far less than the margins).
// Hits are appended to R through the wave's LDS counter.
__device__ __forceinline__ void edge_test(float ax, float ay, float d1x, float d1y, float bx, float by, float d2x,
                                          float d2y, uint32_t ea, uint32_t eb, Rec *R, uint32_t cap, uint32_t *cnt) {
    const float rx = bx - ax, ry = by - ay;
    const float dn = d1x * d2y - d1y * d2x;
    const float c2 = rx * d2y - ry * d2x;
    const float c1 = rx * d1y - ry * d1x;
    const float adn = fabsf(dn);
    const float s2 = dn < 0.f ? -c2 : c2, s1 = dn < 0.f ? -c1 : c1;
    const float lo = -1e-6f * adn, hi = 1.000001f * adn;
    if (!(s2 >= lo && s2 <= hi && s1 >= lo && s1 <= hi && (double)adn >= 1e-12)) return;
    const double denom = dn;
    const double t = c2 / denom;
    const double u = c1 / denom;
    if (t >= -1e-9 && t <= 1 + 1e-9 && u >= -1e-9 && u <= 1 + 1e-9) {
        const float tf = (float)t;
        Rec rec;
        rec.px = ax + tf * d1x;
        rec.py = ay + tf * d1y;
        rec.ei = (int32_t)ea;
        rec.ej = (int32_t)eb;
        rec.t = t;
        rec.u = u;
        const uint32_t pos = atomicAdd(cnt, 1u);
        if (pos < cap) R[pos] = rec;
    }
}

// Edge pairs that passed the line screen below, waiting for the exact test:
// vertex indexes of the source and target edge's first vertex.
constexpr uint32_t kQueue = 128;  // flushed once 64 are queued, so one push always fits
struct EdgeQueue {
    uint32_t ea[kQueue];
    uint32_t eb[kQueue];
};

// The exact test (edge_test) of every queued edge pair, one per lane.
__device__ __forceinline__ void flush_edges(const float *AX, const float *AY, const float *BX, const float *BY,
                                            EdgeQueue &Q, uint32_t &qn, Rec *R, uint32_t cap, uint32_t *cnt,
                                            int lane) {
    wave_sync();
    for (uint32_t k = lane; k < qn; k += 64) {
        const uint32_t ea = Q.ea[k], eb = Q.eb[k];
        const float ax = AX[ea], ay = AY[ea], bx = BX[eb], by = BY[eb];
        edge_test(ax, ay, AX[ea + 1] - ax, AY[ea + 1] - ay, bx, by, BX[eb + 1] - bx, BY[eb + 1] - by, ea, eb, R,
                  cap, cnt);
    }
    qn = 0;
    wave_sync();
}

// Appends the lanes with cand set (ballot prefix); qn is wave-uniform.
__device__ __forceinline__ void push_edge(bool cand, uint32_t ea, uint32_t eb, const float *AX, const float *AY,
                                          const float *BX, const float *BY, EdgeQueue &Q, uint32_t &qn, Rec *R,
                                          uint32_t cap, uint32_t *cnt, int lane) {
    const uint64_t m = __ballot(cand);
    if (m == 0) return;
    if (cand) {
        const uint32_t pos = qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        Q.ea[pos] = ea;
        Q.eb[pos] = eb;
    }
    qn = uni(qn + (uint32_t)__popcll(m));
    if (qn >= 64) flush_edges(AX, AY, BX, BY, Q, qn, R, cap, cnt, lane);
}

// Line screen.  For the source edge a + s*d1 of a lane, o_j = (b_j - a) x d1
// is computed with edge_test's own c1 expression, once per target vertex, and
// serves the two target edges that meet at b_j.  edge_test accepts edge j only
// if u = o_j / dn lies in [-1e-9, 1 + 1e-9]; in exact arithmetic o_{j+1} =
// o_j - dn, and the rounding of o_j, o_{j+1} and dn (each product and
// difference rounded once, f32) moves o_{j+1} - (o_j - dn) by at most
// 2^-21 * |d1|max * (R_j + R_{j+1}), R = |b - a|_1, while 1e-9 * |dn| stays
// below 2^-29 of the same scale.  So an accepted edge has min(o_j, o_{j+1})
// <= m and max(o_j, o_{j+1}) >= -m for m = 2^-19 * |d1|_1 * (R_j + R_{j+1})
// (four times the bound, which absorbs the rounding of m itself); edges that
// fail the screen cannot produce a record, and the rest go to the exact test.
struct LineProbe {
    float o, r;
};
__device__ __forceinline__ LineProbe line_probe(float ax, float ay, float d1x, float d1y, float bx, float by) {
    const float rx = bx - ax, ry = by - ay;
    return {rx * d1y - ry * d1x, fabsf(rx) + fabsf(ry)};
}
__device__ __forceinline__ bool line_pass(LineProbe p0, LineProbe p1, float k) {
    const float m = k * (p0.r + p1.r);
    return (p0.o <= m || p1.o <= m) && (p0.o >= -m || p1.o >= -m);
}

// Target vertices read through the constant address space: the vertex arrays
// do not change during a launch, and uniform addresses then become scalar
// loads.
typedef __attribute__((address_space(4))) const float cfloat;
constexpr int kScreenChunk = 16;  // target vertices per chunk (<= the vertex padding)
static_assert(kScreenChunk == (int)kIxChunk, "the chunk skip reads kIxChunk-segment boxes");

// Queues the lanes of mask (vertex indexes ea per lane, eb uniform).
__device__ __forceinline__ void push_mask(uint64_t mask, uint32_t ea, uint32_t eb, const float *AX, const float *AY,
                                          const float *BX, const float *BY, EdgeQueue &Q, uint32_t &qn, Rec *R,
                                          uint32_t cap, uint32_t *cnt, int lane) {
    if ((mask >> lane) & 1ull) {
        const uint32_t pos = qn + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        Q.ea[pos] = ea;
        Q.eb[pos] = eb;
    }
    qn = uni(qn + (uint32_t)__popcll(mask));
    if (qn >= 64) flush_edges(AX, AY, BX, BY, Q, qn, R, cap, cnt, lane);
}

// One lane's source edge for the chunked screen, and the side masks of a
// target vertex: o = (q - a) x d1 (fused: the screen's own value, within the
// same error bound as the reference's c1, which the margin M covers); bits of
// gt / lt set where o > M / o < -M (fcmp predicates 2 and 4, ordered).
struct ScreenLane {
    float ax, ay, d1x, d1y, M;
};
__device__ __forceinline__ void screen_side(const ScreenLane &L, float qx, float qy, uint64_t &gt, uint64_t &lt) {
    const float o = fmaf(qx - L.ax, L.d1y, -((qy - L.ay) * L.d1x));
    gt = __builtin_amdgcn_fcmpf(o, L.M, 2);
    lt = __builtin_amdgcn_fcmpf(o, -L.M, 4);
}

// Target edges j0 .. j0 + nstep - 1 (nstep == kScreenChunk unless TAIL): the
// vertices of the chunk are uniform scalar loads, which may read up to
// kScreenChunk - 1 vertices past the sequence, inside the zeroed padding that
// follows the vertex arrays at worst.  An edge passes unless both its
// vertices lie beyond the margin on one side.
template <bool TAIL>
__device__ __forceinline__ void screen_chunk(const ScreenLane &L, const cfloat *bxc, const cfloat *byc, uint32_t j0,
                                             uint32_t nstep, uint64_t on_mask, uint64_t &gt, uint64_t &lt,
                                             uint32_t ea, uint32_t lb, const float *AX, const float *AY,
                                             const float *BX, const float *BY, EdgeQueue &Q, uint32_t &qn, Rec *R,
                                             uint32_t cap, uint32_t *cnt, int lane) {
    float qx[kScreenChunk], qy[kScreenChunk];
#pragma unroll
    for (int m = 0; m < kScreenChunk; m++) {
        qx[m] = bxc[j0 + 1 + m];
        qy[m] = byc[j0 + 1 + m];
    }
    // the failing lanes of each step; the active-lane mask is applied once,
    // to the chunk's summary and to the steps that push
    uint64_t fail[kScreenChunk];
    uint64_t all_fail = ~0ull;
#pragma unroll
    for (int m = 0; m < kScreenChunk; m++) {
        uint64_t gt1, lt1;
        screen_side(L, qx[m], qy[m], gt1, lt1);
        fail[m] = (gt & gt1) | (lt & lt1);
        if (TAIL && (uint32_t)m >= nstep) fail[m] = ~0ull;
        all_fail &= fail[m];
        gt = gt1;
        lt = lt1;
    }
    if (~all_fail & on_mask) {
#pragma unroll
        for (int m = 0; m < kScreenChunk; m++) {
            const uint64_t pass = ~fail[m] & on_mask;
            if (pass) push_mask(pass, ea, lb + j0 + m, AX, AY, BX, BY, Q, qn, R, cap, cnt, lane);
        }
    }
}

// All edge pairs of one sequence pair.  Source edges go to lanes and stay in
// registers; when the source sequence fills the wave, the target vertices
// stream in uniform chunks (scalar loads), else lanes split into phases, each
// over a contiguous block of the target edges.  Pairs passing the line screen
// are queued for the exact test.
//
// Chunk skip (cbox: the target sequence's kIxChunk-segment chunk boxes,
// centre and rounded-up half extents; null: none).  Every vertex q of a chunk
// lies in its box, so the exact o(q) = (q - a) x d1 is within
// S = |d1y| hx + |d1x| hy of o at the centre c.  The f32 o of the screen errs
// by at most 2^-22 |d1|_1 R + 2^-24 |o| (R = |q - a|_1 <= Rmax, the box lies
// in the sequence's box), i.e. by M / 16 and a relative ulp, and so does the
// f32 o(c).  When the f32 o(c) exceeds T = S (1 + 2^-20) + 2 M + 2^-100 in
// magnitude, every vertex's f32 o lies beyond M on the side of o(c): the
// screen would fail every edge of the chunk, and the vertices' side masks are
// those of the sign of o(c).  A chunk every active lane skips is passed over
// whole (its vertices are not loaded).  An infinite box (a zero-length
// segment, a non-finite vertex) makes T infinite or NaN and is never skipped.
__device__ __forceinline__ void collect_seq_pair(const float *AX, const float *AY, const float *BX, const float *BY,
                                                 uint2 e1, uint2 e2, const cfloat *cbox, EdgeQueue &Q, uint32_t &qn,
                                                 Rec *R, uint32_t cap, uint32_t *cnt, int lane) {
    const uint32_t la = e1.x, na = e1.y, lb = e2.x, nb = e2.y;  // first edge, edge count
    if (na == 0 || nb == 0) return;
    if (na >= 48) {
        // the target sequence's box bounds every R_j of the screen: with
        // R_j, R_j+1 <= Rmax = max|bx - ax| + max|by - ay| over the box, one
        // margin M = 2^-19 * |d1|_1 * 2 * Rmax >= m serves the whole
        // sequence, and each vertex's side of the line (o > M, o < -M) is a
        // lane mask shared by its two edges
        float bx0 = INFINITY, by0 = INFINITY, nbx1 = INFINITY, nby1 = INFINITY;
        for (uint32_t v = lane; v <= nb; v += 64) {
            const float x = BX[lb + v], y = BY[lb + v];
            bx0 = fminf(bx0, x);
            by0 = fminf(by0, y);
            nbx1 = fminf(nbx1, -x);
            nby1 = fminf(nby1, -y);
        }
        bx0 = wave_min(bx0);
        by0 = wave_min(by0);
        const float bx1 = -wave_min(nbx1), by1 = -wave_min(nby1);
        const cfloat *bxc = (const cfloat *)(BX + lb), *byc = (const cfloat *)(BY + lb);
        for (uint32_t i0 = 0; i0 < na; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool on = i < na;
            const uint32_t ea = la + (on ? i : 0);
            const float ax = AX[ea], ay = AY[ea];
            const float d1x = AX[ea + 1] - ax, d1y = AY[ea + 1] - ay;
            const float rmax = fmaxf(fabsf(bx0 - ax), fabsf(bx1 - ax)) + fmaxf(fabsf(by0 - ay), fabsf(by1 - ay));
            const float M = 0x1p-18f * (fabsf(d1x) + fabsf(d1y)) * rmax;
            const uint64_t on_mask = __ballot(on);
            ScreenLane L{ax, ay, d1x, d1y, M};
            uint64_t gt, lt;
            screen_side(L, bxc[0], byc[0], gt, lt);
            const float ad1x = fabsf(d1x), ad1y = fabsf(d1y);
            const float T0 = fmaf(M, 2.f, 0x1p-100f);
            // true when every active lane skips chunk c (then gt / lt are the
            // side masks of its last vertex)
            auto skip_chunk = [&](uint32_t c) -> bool {
                if (!cbox) return false;
                const float cx = cbox[4 * c], cy = cbox[4 * c + 1], hx = cbox[4 * c + 2], hy = cbox[4 * c + 3];
                const float oc = fmaf(cx - ax, d1y, -((cy - ay) * d1x));
                const float T = fmaf(fmaf(ad1y, hx, ad1x * hy), 1.000001f, T0);
                if (__ballot(on && !(fabsf(oc) > T))) return false;
                gt = __ballot(oc > 0.f);
                lt = __ballot(oc < 0.f);
                return true;
            };
            const uint32_t nfull = nb / kScreenChunk * kScreenChunk;
            for (uint32_t j0 = 0; j0 < nfull; j0 += kScreenChunk) {
                if (skip_chunk(j0 / kScreenChunk)) continue;
                screen_chunk<false>(L, bxc, byc, j0, kScreenChunk, on_mask, gt, lt, ea, lb, AX, AY, BX, BY, Q, qn, R,
                                    cap, cnt, lane);
            }
            if (nfull < nb && !skip_chunk(nfull / kScreenChunk))
                screen_chunk<true>(L, bxc, byc, nfull, nb - nfull, on_mask, gt, lt, ea, lb, AX, AY, BX, BY, Q, qn, R,
                                   cap, cnt, lane);
        }
    } else {
        const uint32_t phases = 64 / na;  // >= 1
        const uint32_t blk = (nb + phases - 1) / phases;
        const uint32_t i = lane % na, ph = lane / na;
        const bool on = ph < phases;
        const uint32_t j0 = on ? min(ph * blk, nb) : nb, j1 = on ? min(j0 + blk, nb) : nb;
        const uint32_t ea = la + i;
        const float ax = AX[ea], ay = AY[ea];
        const float d1x = AX[ea + 1] - ax, d1y = AY[ea + 1] - ay;
        const float k = 0x1p-19f * (fabsf(d1x) + fabsf(d1y));
        LineProbe p0 = line_probe(ax, ay, d1x, d1y, BX[lb + j0], BY[lb + j0]);
        for (uint32_t t = 0; t < blk; t++) {
            const uint32_t j = j0 + t;
            const bool act = j < j1;
            const uint32_t v = act ? j + 1 : j0;  // vertex lb + v is within the sequence
            const LineProbe p1 = line_probe(ax, ay, d1x, d1y, BX[lb + v], BY[lb + v]);
            push_edge(act && line_pass(p0, p1, k), ea, lb + j, AX, AY, BX, BY, Q, qn, R, cap, cnt, lane);
            p0 = p1;
        }
    }
}

// Ideal::intersection for one pair plus the area; R / D: the wave's two
// record slabs of cap records each (cap a power of two).
__device__ __forceinline__ PairRes intersection_pair(const DevSet &S, const double *ring2a, uint32_t ia, uint32_t ib,
                                                     Rec *R, Rec *D, uint32_t cap, uint32_t *cnt, EdgeQueue &Q,
                                                     int lane, const PieceOut &po) {
    PairRes r;
    for (int k = 0; k < 5; k++) r.c[k] = 0;
    r.area = 0.0;
    r.bb = r.tests = r.raw = 0;
    r.over = false;
    const farm_raster_hdr &hA = S.hdr[ia];
    const farm_raster_hdr &hB = S.hdr[ib];
    // 1. source pixels over the target MBB (retrieve_pixels, +-1e-6)
    const Window w = window_of(hA, hB.mbr[0], hB.mbr[1], hB.mbr[2], hB.mbr[3]);
    const int wx = w.x1 - w.x0 + 1, wy = w.y1 - w.y0 + 1;
    const uint32_t npx = (wx > 0 && wy > 0) ? (uint32_t)(wx * wy) : 0u;
    uint32_t etn = 0, itn = 0;
    for (uint32_t c0 = 0; c0 < npx; c0 += 64) {
        const uint32_t k = c0 + lane;
        uint8_t st = 1;
        if (k < npx) {
            const int x = w.x0 + (int)(k / wy), y = w.y0 + (int)(k % wy);
            st = S.status[hA.pix_start + y * hA.dimx + x];
        }
        etn += __popcll(__ballot(k < npx && st == 0));
        itn += __popcll(__ballot(k < npx && st == kCC - 1));
    }
    if (etn == npx) {
        r.c[0] = -1;
        return r;
    }
    if (itn == npx) {
        r.c[0] = -2;
        r.area = fabs(ring2a[ib]) / 2.0;
        put_pieces(po, lane == 0, farm_piece{po.pair, 1, 0, -1, 0.f, 0.f, 0.f, 0.f}, lane);  // all of B
        return r;
    }
    // 2. crossing points of the edges of BORDER x BORDER pixel pairs
    const float *AX = S.x + hA.vert_start, *AY = S.y + hA.vert_start;
    const float *BX = S.x + hB.vert_start, *BY = S.y + hB.vert_start;
    if (lane == 0) *cnt = 0u;
    wave_sync();
    uint32_t qn = 0;
    for (int x = w.x0; x <= w.x1; x++) {
        for (int y = w.y0; y <= w.y1; y++) {
            const uint32_t pa = (uint32_t)(y * hA.dimx + x);
            if (show(S.status[hA.pix_start + pa]) != 1) continue;
            double lx, ly, hx, hy;
            pixel_box(hA, x, y, lx, ly, hx, hy);
            const Window v = window_of(hB, lx, ly, hx, hy);
            for (int x2 = v.x0; x2 <= v.x1; x2++) {
                for (int y2 = v.y0; y2 <= v.y1; y2++) {
                    const uint32_t pb = (uint32_t)(y2 * hB.dimx + x2);
                    if (show(S.status[hB.pix_start + pb]) != 1) continue;
                    r.bb++;
                    const uint32_t sa0 = S.pixoff[hA.pixoff_start + pa], sa1 = S.pixoff[hA.pixoff_start + pa + 1];
                    const uint32_t sb0 = S.pixoff[hB.pixoff_start + pb], sb1 = S.pixoff[hB.pixoff_start + pb + 1];
                    for (uint32_t sa = sa0; sa < sa1; sa++) {
                        const uint2 e1 = S.seqs[hA.seq_start + sa];
                        for (uint32_t sb = sb0; sb < sb1; sb++) {
                            const uint2 e2 = S.seqs[hB.seq_start + sb];
                            r.tests += (uint64_t)e1.y * e2.y;
                            const cfloat *cb =
                                S.ixbox ? (const cfloat *)(S.ixbox + S.ixoff[hB.seq_start + sb]) : nullptr;
                            collect_seq_pair(AX, AY, BX, BY, e1, e2, cb, Q, qn, R, cap, cnt, lane);
                        }
                    }
                }
            }
        }
    }
    if (qn) flush_edges(AX, AY, BX, BY, Q, qn, R, cap, cnt, lane);
    wave_sync();
    const uint32_t nrec = uni(*cnt);
    r.raw = nrec;
    if (nrec > cap) {
        r.over = true;
        return r;
    }
    wave_sync();
    const double ox = fmax(hA.mbr[0], hB.mbr[0]), oy = fmax(hA.mbr[1], hB.mbr[1]);
    uint32_t nk = 0;
    if (nrec > 0) {
        // 3. sort by point, keep the first of each run of equal points
        bitonic<0>(R, nrec, lane);
        bool small = false;
        for (uint32_t k = lane; k < nrec; k += 64) {
            const Rec e = R[k];
            small |= fabsf(e.px) < 0.015625f || fabsf(e.py) < 0.015625f;
        }
        if (!wave_any(small)) {
            // above 2^-6 the 1e-9 tolerance is float equality: a record is
            // kept when it differs from its predecessor
            for (uint32_t c0 = 0; c0 < nrec; c0 += 64) {
                const uint32_t k = c0 + lane;
                bool keep = false;
                Rec e;
                if (k < nrec) {
                    e = R[k];
                    keep = k == 0 || !pt_eq(R[k - 1].px, R[k - 1].py, e.px, e.py);
                }
                const uint64_t km = __ballot(keep);
                if (keep) D[nk + (uint32_t)__popcll(km & ((1ull << lane) - 1ull))] = e;
                nk += (uint32_t)__popcll(km);
            }
        } else if (lane == 0) {
            // std::unique: compare with the last kept record
            for (uint32_t k = 0; k < nrec; k++) {
                const Rec e = R[k];
                if (nk == 0 || !pt_eq(D[nk - 1].px, D[nk - 1].py, e.px, e.py)) D[nk++] = e;
            }
        }
        nk = uni(__shfl(nk, 0, 64));
        wave_sync();
    }
    r.c[0] = (int32_t)nk;
    uint32_t pieces = 0;
    double acc_a = 0.0, acc_b = 0.0;
    if (nk > 0) {
        bitonic<1>(D, nk, lane);
        acc_a = ring_pieces<true>(S, hA, hB, D, nk, ox, oy, lane, r, pieces, po);
        wave_sync();
        bitonic<2>(D, nk, lane);
        acc_b = ring_pieces<false>(S, hB, hA, D, nk, ox, oy, lane, r, pieces, po);
    }
    if (pieces == 0) {
        // containment or disjoint
        if (seg_contain_wave(S, hB, AX[0], AY[0], lane) == 2) {
            r.area = fabs(ring2a[ia]) / 2.0;
            put_pieces(po, lane == 0, farm_piece{po.pair, 0, 0, -1, 0.f, 0.f, 0.f, 0.f}, lane);  // all of A
        } else if (seg_contain_wave(S, hA, BX[0], BY[0], lane) == 2) {
            r.area = fabs(ring2a[ib]) / 2.0;
            put_pieces(po, lane == 0, farm_piece{po.pair, 1, 0, -1, 0.f, 0.f, 0.f, 0.f}, lane);  // all of B
        }
        return r;
    }
    const double sa = ring2a[ia] < 0 ? -1.0 : 1.0, sb = ring2a[ib] < 0 ? -1.0 : 1.0;
    const double side_a = wave_sum(acc_a), side_b = wave_sum(acc_b);
    r.area = fabs(sa * side_a + sb * side_b) / 2.0;
    return r;
}

}  // namespace

// Twice the signed area of every ring, relative to its MBB's low corner
// (one wave per polygon).
__global__ __launch_bounds__(256) void k_ring_area(DevSet S, int64_t npolys, double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= npolys) return;
    const farm_raster_hdr &h = S.hdr[p];
    const float *X = S.x + h.vert_start, *Y = S.y + h.vert_start;
    const double ox = h.mbr[0], oy = h.mbr[1];
    double acc = 0.0;
    for (uint32_t k = lane; k + 1 < h.nverts; k += 64)
        acc += ((double)X[k] - ox) * ((double)Y[k + 1] - oy) - ((double)X[k + 1] - ox) * ((double)Y[k] - oy);
    acc = wave_sum(acc);
    if (lane == 0) out[p] = acc;
}

// Persistent waves over pairs; BIG: the pairs in `list` with larger slabs.
template <bool BIG>
__global__ __launch_bounds__(256) void k_intersection(DevSet S, const double *__restrict__ ring2a,
                                                      const uint32_t *__restrict__ pa, const uint32_t *__restrict__ pb,
                                                      uint32_t n, const uint32_t *__restrict__ list,
                                                      int32_t *__restrict__ counts, double *__restrict__ area,
                                                      Rec *__restrict__ slab, uint32_t cap,
                                                      uint32_t *__restrict__ ctr, uint32_t *__restrict__ ovf,
                                                      unsigned long long *__restrict__ stats,
                                                      farm_piece *__restrict__ pcs, uint32_t pcap) {
    const int lane = threadIdx.x & 63;
    const uint64_t gw = (uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    Rec *R = slab + gw * 2ull * cap;
    Rec *D = R + cap;
    __shared__ uint32_t rec_cnt[kWaves];
    __shared__ EdgeQueue queues[kWaves];
    uint32_t *cnt = &rec_cnt[threadIdx.x >> 6];
    EdgeQueue &Q = queues[threadIdx.x >> 6];
    unsigned long long bb = 0, tests = 0, raw = 0, found = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&ctr[0], kChunk);
        base = uni(__shfl(base, 0, 64));
        if (base >= n) break;
        const uint32_t end = min(base + kChunk, n);
        for (uint32_t k = base; k < end; k++) {
            const uint32_t idx = BIG ? list[k] : k;
            const uint32_t ia = pa[idx], ib = pb[idx];
            if ((S.hdr[ia].flags | S.hdr[ib].flags) & FARM_HDR_REJECTED) {
                if (lane < 5) counts[5ull * idx + lane] = lane == 0 ? -3 : 0;
                if (lane == 0) area[idx] = 0.0;
                continue;
            }
            const PieceOut po{pcs, ctr + 2, pcap, idx};
            const PairRes r = intersection_pair(S, ring2a, ia, ib, R, D, cap, cnt, Q, lane, po);
            if (r.over) {
                if (lane == 0) {
                    if (BIG) {  // still too many: the host reruns it with larger slabs
                        counts[5ull * idx] = -4;
                        counts[5ull * idx + 1] = (int32_t)min(r.raw, (uint32_t)INT_MAX);
                    } else {
                        const uint32_t slot = atomicAdd(&ctr[1], 1u);
                        ovf[slot] = idx;
                    }
                }
                continue;
            }
            bb += r.bb;
            tests += r.tests;
            raw += r.raw;
            found += r.area > 0.0;
            if (lane < 5) counts[5ull * idx + lane] = r.c[lane];
            if (lane == 0) area[idx] = r.area;
        }
    }
    if (lane == 0) {
        atomicAdd(&stats[0], bb);
        atomicAdd(&stats[1], tests);
        atomicAdd(&stats[2], raw);
        atomicAdd(&stats[3], found);
    }
}

template __global__ void k_intersection<false>(DevSet, const double *, const uint32_t *, const uint32_t *, uint32_t,
                                               const uint32_t *, int32_t *, double *, Rec *, uint32_t, uint32_t *,
                                               uint32_t *, unsigned long long *, farm_piece *, uint32_t);
template __global__ void k_intersection<true>(DevSet, const double *, const uint32_t *, const uint32_t *, uint32_t,
                                              const uint32_t *, int32_t *, double *, Rec *, uint32_t, uint32_t *,
                                              uint32_t *, unsigned long long *, farm_piece *, uint32_t);

}  // namespace farm

using namespace farm;

namespace farm {
int build_ix_boxes(farm_handle h);
}  // namespace farm

namespace {

int run_intersection(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts_dev, double *area_dev,
                     farm_stats *st, farm_piece *pcs_dev = nullptr, uint32_t pcap = 0, uint64_t *npcs = nullptr) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    if (h->has_layers) return fail(h, FARM_ESTATE, "intersection needs a set rasterized without -h");
    if (begin > end || end > h->npairs)
        return fail(h, FARM_EINVAL, "pair window [%llu,%llu) outside [0,%llu)", (unsigned long long)begin,
                    (unsigned long long)end, (unsigned long long)h->npairs);
    const uint64_t n64 = end - begin;
    if (n64 > 0xffffffffull) return fail(h, FARM_ECAPACITY, "batch larger than 2^32 pairs");
    const uint32_t n = (uint32_t)n64;
    HIPCHK(h, hipSetDevice(h->device));
    if (st) memset(st, 0, sizeof *st);
    if (st) st->pairs = n;
    if (n == 0) return FARM_OK;
    int rc;
    if ((rc = build_ix_boxes(h))) return rc;
    DevSet S = devset(h);
    if (!h->ring_area_ready) {
        if ((rc = ensure(h, h->ring2a, 8ull * h->npolys))) return rc;
        hipLaunchKernelGGL(k_ring_area, dim3((unsigned)((h->npolys + 3) / 4)), dim3(256), 0, h->stream, S, h->npolys,
                           P<double>(h->ring2a));
        HIPCHK(h, hipGetLastError());
        h->ring_area_ready = true;
    }
    const uint32_t *pa = P<uint32_t>(h->pa) + begin;
    const uint32_t *pb = P<uint32_t>(h->pb) + begin;
    if ((rc = ensure(h, h->wctr, 64))) return rc;
    if ((rc = ensure(h, h->wovf, 4ull * n))) return rc;
    uint32_t *ctr = P<uint32_t>(h->wctr);
    unsigned long long *ist = reinterpret_cast<unsigned long long *>(ctr + 8);
    if ((rc = fill_u32(h, ctr, 16, 0))) return rc;
    const unsigned nblocks = (unsigned)h->num_cus * 4;
    const uint64_t nwaves = (uint64_t)nblocks * kWaves;
    if ((rc = ensure(h, h->wslab, nwaves * 2ull * kRecCap * sizeof(Rec)))) return rc;
    HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    hipLaunchKernelGGL(k_intersection<false>, dim3(nblocks), dim3(64 * kWaves), 0, h->stream, S,
                       (const double *)P<double>(h->ring2a), pa, pb, n, (const uint32_t *)nullptr, counts_dev, area_dev,
                       P<Rec>(h->wslab), kRecCap, ctr, P<uint32_t>(h->wovf), ist, pcs_dev, pcap);
    HIPCHK(h, hipGetLastError());
    uint32_t novf = 0;
    HIPCHK(h, hipMemcpyAsync(&novf, ctr + 1, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (novf) {
        // pairs with more raw crossing records than the first slabs hold
        const unsigned bblocks = 16;
        if ((rc = ensure(h, h->wslab_big, (uint64_t)bblocks * kWaves * 2ull * kRecCapBig * sizeof(Rec)))) return rc;
        if ((rc = fill_u32(h, ctr, 1, 0))) return rc;
        hipLaunchKernelGGL(k_intersection<true>, dim3(bblocks), dim3(64 * kWaves), 0, h->stream, S,
                           (const double *)P<double>(h->ring2a), pa, pb, novf, (const uint32_t *)P<uint32_t>(h->wovf),
                           counts_dev, area_dev, P<Rec>(h->wslab_big), kRecCapBig, ctr, P<uint32_t>(h->wovf), ist,
                           pcs_dev, pcap);
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    unsigned long long hst[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(hst, ist, sizeof hst, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (novf) {
        // pairs beyond the second slabs: once more, with slabs sized to the
        // largest of them (fewer waves, at most 8 GiB of slabs)
        std::vector<uint32_t> list(novf), big;
        HIPCHK(h, hipMemcpy(list.data(), P<uint32_t>(h->wovf), 4ull * novf, hipMemcpyDeviceToHost));
        uint32_t most = 0;
        for (uint32_t idx : list) {
            int32_t c[2];
            HIPCHK(h, hipMemcpy(c, counts_dev + 5ull * idx, 8, hipMemcpyDeviceToHost));
            if (c[0] == -4) {
                big.push_back(idx);
                most = std::max(most, (uint32_t)c[1]);
            }
        }
        if (!big.empty()) {
            uint64_t cap = 2ull * kRecCapBig;
            while (cap < most) cap <<= 1;
            if (cap > (1ull << 26))
                return fail(h, FARM_ECAPACITY, "a pair has %u crossing records (limit 2^26)", most);
            const uint64_t per_wave = 2ull * cap * sizeof(Rec);
            const uint64_t waves = std::max<uint64_t>(kWaves, std::min<uint64_t>(64, (8ull << 30) / per_wave));
            const unsigned blocks = (unsigned)(waves / kWaves);
            if ((rc = ensure(h, h->wslab_big, (uint64_t)blocks * kWaves * per_wave))) return rc;
            HIPCHK(h, hipMemcpyAsync(P<uint32_t>(h->wovf), big.data(), 4ull * big.size(), hipMemcpyHostToDevice,
                                     h->stream));
            if ((rc = fill_u32(h, ctr, 1, 0))) return rc;
            hipLaunchKernelGGL(k_intersection<true>, dim3(blocks), dim3(64 * kWaves), 0, h->stream, S,
                               (const double *)P<double>(h->ring2a), pa, pb, (uint32_t)big.size(),
                               (const uint32_t *)P<uint32_t>(h->wovf), counts_dev, area_dev, P<Rec>(h->wslab_big),
                               (uint32_t)cap, ctr, P<uint32_t>(h->wovf), ist, pcs_dev, pcap);
            HIPCHK(h, hipGetLastError());
            HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
            HIPCHK(h, hipMemcpyAsync(hst, ist, sizeof hst, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (uint32_t idx : big) {
                int32_t c0 = 0;
                HIPCHK(h, hipMemcpy(&c0, counts_dev + 5ull * idx, 4, hipMemcpyDeviceToHost));
                if (c0 == -4)
                    return fail(h, FARM_ECAPACITY, "pair %llu: more than %llu crossing records",
                                (unsigned long long)(begin + idx), (unsigned long long)cap);
            }
        }
    }
    if (npcs) {
        uint32_t np = 0;
        HIPCHK(h, hipMemcpy(&np, ctr + 2, 4, hipMemcpyDeviceToHost));
        *npcs = np;
    }
    if (st) {
        float ms = 0;
        HIPCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        st->bb_pixel_pairs = hst[0];
        st->segment_tests = hst[1];
        st->refine_items = hst[2];
        st->found = hst[3];
        st->refine_ms = ms;
        st->total_ms = ms;
    }
    return FARM_OK;
}

}  // namespace

extern "C" {

int farm_intersection_area_dev(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts_dev, double *area_dev,
                               farm_stats *st) {
    if (!counts_dev || !area_dev) return fail(h, FARM_EINVAL, "null output buffer");
    return run_intersection(h, begin, end, counts_dev, area_dev, st);
}

int farm_intersection_pieces(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts, double *area,
                             farm_piece *pieces, uint64_t cap, uint64_t *npieces, farm_stats *st) {
    if (!h) return FARM_EINVAL;
    if (begin > end || !npieces || (cap && !pieces)) return fail(h, FARM_EINVAL, "bad arguments");
    if (cap > 0xffffffffull) cap = 0xffffffffull;
    const uint64_t n = end - begin;
    int rc;
    if ((rc = ensure(h, h->icounts, 20 * n + 16))) return rc;
    if ((rc = ensure(h, h->iarea, 8 * n + 16))) return rc;
    if ((rc = ensure(h, h->ipieces, sizeof(farm_piece) * (cap ? cap : 1)))) return rc;
    if ((rc = run_intersection(h, begin, end, P<int32_t>(h->icounts), P<double>(h->iarea), st,
                               P<farm_piece>(h->ipieces), (uint32_t)cap, npieces)))
        return rc;
    if (n && counts) HIPCHK(h, hipMemcpy(counts, h->icounts.p, 20 * n, hipMemcpyDeviceToHost));
    if (n && area) HIPCHK(h, hipMemcpy(area, h->iarea.p, 8 * n, hipMemcpyDeviceToHost));
    if (*npieces > cap)
        return fail(h, FARM_ECAPACITY, "%llu pieces, room for %llu", (unsigned long long)*npieces,
                    (unsigned long long)cap);
    if (*npieces) HIPCHK(h, hipMemcpy(pieces, h->ipieces.p, sizeof(farm_piece) * *npieces, hipMemcpyDeviceToHost));
    return FARM_OK;
}

int farm_intersection_area(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts, double *area,
                           farm_stats *st) {
    if (!h) return FARM_EINVAL;
    if (begin > end) return fail(h, FARM_EINVAL, "bad pair window");
    const uint64_t n = end - begin;
    int rc;
    if ((rc = ensure(h, h->icounts, 20 * n + 16))) return rc;
    if ((rc = ensure(h, h->iarea, 8 * n + 16))) return rc;
    if ((rc = run_intersection(h, begin, end, P<int32_t>(h->icounts), P<double>(h->iarea), st))) return rc;
    if (n && counts) HIPCHK(h, hipMemcpy(counts, h->icounts.p, 20 * n, hipMemcpyDeviceToHost));
    if (n && area) HIPCHK(h, hipMemcpy(area, h->iarea.p, 8 * n, hipMemcpyDeviceToHost));
    return FARM_OK;
}

}  // extern "C"
