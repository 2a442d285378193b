// The IDEAL raster model of a whole polygon set, built on the device.  The
// bytes equal the host builder's (host/raster.cpp), which restates the
// reference CPU builder step for step:
//   grid sizing        MyRaster::init_raster(int)       MyRaster.cpp:8-33 (host, per polygon)
//   edge walk          Ideal::evaluate_edges            Ideal.cpp:252-538
//   edge sequences     Ideal::process_crosses           Ideal.cpp:90-183
//   grid-line nodes    Ideal::process_intersection      Ideal.cpp:185-238
//   CSR fill-in        Ideal::process_pixels_null       Ideal.cpp:35-45
//   IN/OUT             Ideal::scanline_reandering       Ideal.cpp:540-573
//   fill classes       Ideal::calculate_fullness        Ideal.cpp:820-918
//   hierarchy          MyRaster::grid_align (host), Ideal::layering (host),
//                      Ideal::merge_status              MyRaster.cpp:444-491, Ideal.cpp:1008-1099
//
// How the work is laid out:
//   - one workgroup per polygon walks its edges 256 at a time, once to
//     count and once to write; block scans give every edge its output
//     positions, so events and grid-line nodes keep the ring order;
//   - a stable radix sort by global pixel id gathers each pixel's events in
//     that order, and one lane per pixel builds the pixel's edge sequences;
//   - grid-line nodes: two stable radix sorts (by value, then by line);
//   - scanline parity: one lane per raster row;
//   - fill classes: one workgroup per polygon clips the ring to each column
//     strip.  Points are deduplicated first-come by a hash of their bit
//     patterns; points within 0.0162 of an axis, where the reference's
//     tolerant equality is not bit equality, take that comparison against
//     every kept point in order.  Then one lane per BORDER pixel clips its
//     column strip to its row and sums the shoelace in the host's order;
//   - one lane per coarse pixel merges the hierarchy;
//   - rejected polygons are emptied and every array is packed.
// Every f64 and f32 operation that reaches a byte is the host's, in the
// same order; conversions to int follow x86-64 (out of range -> INT_MIN).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <chrono>
#include <hipcub/hipcub.hpp>
#include <thread>
#include <vector>

#include "../../../include/farm_host.h"
#include "../host/raster.h"
#include "handle.h"
#include "wave.h"

namespace farm {

namespace {

constexpr uint8_t kRawBorder = 1;
constexpr uint8_t kRawIn = 2;
constexpr int kRB = 256;              // threads of the per-polygon workgroups
constexpr uint32_t kLdsCands = 1024;  // strips up to this many candidates hash in LDS
constexpr uint32_t kLdsSlots = 2048;
constexpr uint32_t kChunkCands = 4 * kRB;
// |v| >= kNearAxis on both axes: the reference's |dx| < 1e-9 equality of
// float points is bit equality (the float spacing there is >= 2^-29)
constexpr float kNearAxis = 0.0162f;

// rejection causes (any makes the polygon an empty raster)
constexpr uint32_t kRejAlign = 1, kRejWalk = 2, kRejSeq = 4, kRejFill = 8, kRejUnsupported = 16;

// Per polygon, from the host grid sizing.
struct RPoly {
    double x0, y0, sx, sy;  // grid origin and pixel steps
    int32_t dimx, dimy;
    uint64_t pix;  // first pixel (global id)
    uint64_t vl;   // first vertical line slot (dimx + 2 per polygon)
    uint64_t hl;   // first horizontal line slot (dimy + 2 per polygon)
    uint64_t v;    // first vertex
    uint32_t nv;   // ring length incl. the repeated first vertex
    uint32_t pad;
};

// (int)v as x86-64 converts (cvttsd2si): NaN and out-of-range give INT_MIN
__device__ __forceinline__ int cvt_x86(double v) {
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

// util.h:98-106 double_to_int with the x86 conversion (i + 1 wraps)
__device__ __forceinline__ int dbl2int_x86(double v) {
    int i = cvt_x86(v);
    const int ip1 = (int)((unsigned)i + 1u);
    if (fabs(1.0 * ip1 - v) < 0.00000001) i = ip1;
    return i;
}

// std::min / std::max on floats (the first argument wins ties and NaNs)
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ bool within(double lo, double hi, double v) { return !(v < lo || v > hi); }

// Point::operator== (Point.h:32-33)
__device__ __forceinline__ bool same_pt(float2 a, float2 b) {
    const float dx = a.x - b.x, dy = a.y - b.y;
    return fabsf(dx) < 1e-9 && fabsf(dy) < 1e-9;
}

__device__ __forceinline__ bool near_axis(float2 p) { return fabsf(p.x) < kNearAxis || fabsf(p.y) < kNearAxis; }

// Exclusive prefix sum over a 256-thread block; sh holds 4 words.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t wt;
    const uint32_t ex = wave_excl_scan(v, lane, &wt);
    if (lane == 0) sh[w] = wt;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kRB / 64; k++) {
        const uint32_t t = sh[k];
        if (k < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + ex;
}

// ------------------------------------------------------------- MBB ----

// VertexSequence::getMBR (vertex_sequence.cpp:152-167) and the largest |x|
// of every ring (the x-strip bucket margin's scale).
__global__ __launch_bounds__(kRB) void k_rmbb(const float *__restrict__ X, const float *__restrict__ Y,
                                              const uint64_t *__restrict__ voff, double *__restrict__ mbb,
                                              float *__restrict__ absx) {
    const int64_t i = blockIdx.x;
    const uint64_t b = voff[i], e = voff[i + 1];
    double lx = 100000.0, ly = 100000.0, hx = -100000.0, hy = -100000.0;
    float ax = 0.f;
    for (uint64_t k = b + threadIdx.x; k < e; k += kRB) {
        const double vx = X[k], vy = Y[k];
        lx = (vx < lx) ? vx : lx;
        hx = (hx < vx) ? vx : hx;
        ly = (vy < ly) ? vy : ly;
        hy = (hy < vy) ? vy : hy;
        ax = smax(ax, fabsf(X[k]));
    }
    __shared__ double s[4][kRB];
    __shared__ float sa[kRB];
    s[0][threadIdx.x] = lx;
    s[1][threadIdx.x] = ly;
    s[2][threadIdx.x] = hx;
    s[3][threadIdx.x] = hy;
    sa[threadIdx.x] = ax;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 1; t < kRB; t++) {
            lx = (s[0][t] < lx) ? s[0][t] : lx;
            ly = (s[1][t] < ly) ? s[1][t] : ly;
            hx = (hx < s[2][t]) ? s[2][t] : hx;
            hy = (hy < s[3][t]) ? s[3][t] : hy;
            ax = smax(ax, sa[t]);
        }
        const float bias = 1e-9f;
        mbb[4 * i + 0] = lx - bias;
        mbb[4 * i + 1] = ly - bias;
        mbb[4 * i + 2] = hx + bias;
        mbb[4 * i + 3] = hy + bias;
        absx[i] = ax;
    }
}

// Pixel -> polygon map and zeroed per-pixel arrays.
__global__ __launch_bounds__(kRB) void k_rinit(const RPoly *__restrict__ P, uint32_t *__restrict__ pixpoly) {
    const int64_t i = blockIdx.x;
    const RPoly g = P[i];
    const uint64_t n = (uint64_t)g.dimx * g.dimy;
    for (uint64_t p = threadIdx.x; p < n; p += kRB) pixpoly[g.pix + p] = (uint32_t)i;
}

// ------------------------------------------------------- edge walk ----

// One edge of Ideal::evaluate_edges: BORDER marks, LEAVE/ENTER events and
// grid-line nodes through the sink; s.ok = false where the reference
// asserts or never terminates.
template <class Sink>
__device__ void walk_edge(const RPoly &g, int e, float fax, float fay, float fbx, float fby, Sink &s) {
    const int dimx = g.dimx, dimy = g.dimy;
    const double x0 = g.x0, y0 = g.y0, sx = g.sx, sy = g.sy;
    const double ax = fax, ay = fay, bx = fbx, by = fby;
    int cx0 = dbl2int_x86((ax - x0) / sx), cx1 = dbl2int_x86((bx - x0) / sx);
    int cy0 = dbl2int_x86((ay - y0) / sy), cy1 = dbl2int_x86((by - y0) / sy);
    if (cx0 == dimx) cx0--;
    if (cx1 == dimx) cx1--;
    if (cy0 == dimy) cy0--;
    if (cy1 == dimy) cy1--;
    if (cx0 >= dimx || cx1 >= dimx || cy0 >= dimy || cy1 >= dimy || cx0 < 0 || cx1 < 0 || cy0 < 0 || cy1 < 0) {
        s.ok = false;
        return;
    }
    s.border(cy0 * dimx + cx0);
    s.border(cy1 * dimx + cx1);
    if (cx0 == cx1 && cy0 == cy1) return;
    // one pixel-to-pixel step: LEAVE the old pixel, ENTER the new one
    auto step = [&](int ox, int oy, int nx, int ny) -> bool {
        if (ox < 0 || ox >= dimx || oy < 0 || oy >= dimy || nx < 0 || nx >= dimx || ny < 0 || ny >= dimy) {
            s.ok = false;
            return false;
        }
        const int a = oy * dimx + ox, b = ny * dimx + nx;
        s.border(a);
        s.event(a, e, 1);
        s.event(b, e, 0);
        s.border(b);
        return true;
    };
    if (ay == by) {  // horizontal edge
        const int d = cx0 < cx1 ? 1 : -1;
        for (int x = cx0; x != cx1; x += d) {
            s.vnode(d > 0 ? x + 1 : x, ay);
            if (!step(x, cy0, x + d, cy0)) return;
        }
        return;
    }
    if (ax == bx) {  // vertical edge
        const int d = cy0 < cy1 ? 1 : -1;
        for (int y = cy0; y != cy1; y += d) {
            s.hnode(d > 0 ? y + 1 : y, ax);
            if (!step(cx0, y, cx0, y + d)) return;
        }
        return;
    }
    // general edge: follow y = k x + c through the grid
    const double k = (ay - by) / (ax - bx);
    const double c = (ax * by - bx * ay) / (ax - bx);
    const int dx = cx0 < cx1 ? 1 : -1, dy = cy0 < cy1 ? 1 : -1;
    const int ylo = min(cy0, cy1), yhi = max(cy0, cy1);
    const int xlo = min(cx0, cx1), xhi = max(cx0, cx1);
    int x = cx0, y = cy0;
    int64_t guard = 4LL * ((int64_t)dimx + dimy) + 16;
    while (x != cx1 || y != cy1) {
        if (guard-- <= 0) {
            s.ok = false;
            return;
        }
        bool moved = false;
        if (x != cx1) {
            const double xv = dx > 0 ? ((double)x + 1) * sx + x0 : (double)x * sx + x0;
            const double yv = xv * k + c;
            const double q = (yv - y0) / sy;
            if (fabs(q - round(q)) < 1e-9) {
                // through a grid corner: a diagonal step (the rightward walks
                // record a vertical node); an edge within one pixel row has
                // no matching case in the reference and its walk never ends
                if (cy0 == cy1) {
                    s.ok = false;
                    return;
                }
                if (dx > 0) s.vnode(dy > 0 ? x + 1 : x, yv);
                if (!step(x, y, x + dx, y + dy)) return;
                x += dx;
                y += dy;
                continue;
            }
            int ry = cvt_x86((yv - y0) / sy);
            ry = min(max(ry, ylo), yhi);
            if (ry == y) {
                moved = true;
                s.vnode(dx > 0 ? x + 1 : x, yv);
                if (!step(x, y, x + dx, y)) return;
                x += dx;
            }
        }
        if (y != cy1) {
            const double yv = dy > 0 ? (double)(y + 1) * sy + y0 : (double)y * sy + y0;
            const double xv = (yv - c) / k;
            int rx = cvt_x86((xv - x0) / sx);
            rx = min(max(rx, xlo), xhi);
            if (rx == x) {
                moved = true;
                s.hnode(dy > 0 ? y + 1 : y, xv);
                if (!step(x, y, x, y + dy)) return;
                y += dy;
            }
        }
        if (!moved) {
            s.ok = false;
            return;
        }
    }
}

struct CountSink {
    bool ok = true;
    uint32_t nev = 0, nh = 0, nvn = 0;
    uint8_t *st = nullptr;  // BORDER marks (count pass only)
    __device__ void border(int p) {
        if (st) st[p] = kRawBorder;
    }
    __device__ void event(int, int, int) { nev++; }
    __device__ void vnode(int, double) { nvn++; }
    __device__ void hnode(int, double) { nh++; }
};

struct WriteSink {
    bool ok = true;
    const RPoly *g;
    uint64_t ev, h, v;  // next output positions
    uint32_t *evkey, *evval, *vnkey, *hnkey, *pixev, *vlc, *hlc;
    double *vnval, *hnval;
    __device__ void border(int) {}
    __device__ void event(int p, int e, int leave) {
        const uint32_t gp = (uint32_t)(g->pix + (uint64_t)p);
        evkey[ev] = gp;
        evval[ev] = ((uint32_t)e << 1) | (uint32_t)leave;
        ev++;
        atomicAdd(&pixev[gp], 1u);
    }
    __device__ void vnode(int line, double val) {
        if (line < 0 || line > g->dimx + 1) {
            ok = false;
            return;
        }
        const uint32_t l = (uint32_t)(g->vl + (uint64_t)line);
        vnkey[v] = l;
        vnval[v] = val;
        v++;
        atomicAdd(&vlc[l], 1u);
    }
    __device__ void hnode(int line, double val) {
        if (line < 0 || line > g->dimy + 1) {
            ok = false;
            return;
        }
        const uint32_t l = (uint32_t)(g->hl + (uint64_t)line);
        hnkey[h] = l;
        hnval[h] = val;
        h++;
        atomicAdd(&hlc[l], 1u);
    }
};

// Count pass: per polygon events / nodes, BORDER marks, walk failures.
__global__ __launch_bounds__(kRB) void k_rwalk_count(const RPoly *__restrict__ P, const float *__restrict__ X,
                                                     const float *__restrict__ Y, uint8_t *__restrict__ status,
                                                     uint32_t *__restrict__ rflag, uint64_t *__restrict__ evc,
                                                     uint64_t *__restrict__ hnc, uint64_t *__restrict__ vnc) {
    const int64_t i = blockIdx.x;
    __shared__ uint32_t sh[4];
    if (rflag[i]) {
        if (threadIdx.x == 0) evc[i] = hnc[i] = vnc[i] = 0;
        return;
    }
    const RPoly g = P[i];
    const int ne = (int)g.nv - 1;
    uint64_t nev = 0, nh = 0, nvn = 0;
    bool ok = true;
    for (int e = threadIdx.x; e < ne; e += kRB) {
        CountSink s;
        s.st = status + g.pix;
        walk_edge(g, e, X[g.v + e], Y[g.v + e], X[g.v + e + 1], Y[g.v + e + 1], s);
        ok &= s.ok;
        nev += s.nev;
        nh += s.nh;
        nvn += s.nvn;
    }
    const bool bad = __syncthreads_or(!ok);
    // the totals, per wave then per block
    __shared__ unsigned long long tot[3];
    if (threadIdx.x == 0) tot[0] = tot[1] = tot[2] = 0;
    __syncthreads();
    atomicAdd(&tot[0], (unsigned long long)nev);
    atomicAdd(&tot[1], (unsigned long long)nh);
    atomicAdd(&tot[2], (unsigned long long)nvn);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (bad) atomicOr(&rflag[i], kRejWalk);
        evc[i] = bad ? 0 : tot[0];
        hnc[i] = bad ? 0 : tot[1];
        vnc[i] = bad ? 0 : tot[2];
    }
    (void)sh;
}

// Write pass: events and nodes at their ring-order positions.
__global__ __launch_bounds__(kRB) void k_rwalk_write(const RPoly *__restrict__ P, const float *__restrict__ X,
                                                     const float *__restrict__ Y, uint32_t *__restrict__ rflag,
                                                     const uint64_t *__restrict__ evo, const uint64_t *__restrict__ hno,
                                                     const uint64_t *__restrict__ vno, uint32_t *evkey,
                                                     uint32_t *evval, uint32_t *vnkey, double *vnval,
                                                     uint32_t *hnkey, double *hnval, uint32_t *pixev, uint32_t *vlc,
                                                     uint32_t *hlc) {
    const int64_t i = blockIdx.x;
    if (rflag[i]) return;
    const RPoly g = P[i];
    const int ne = (int)g.nv - 1;
    __shared__ uint32_t sh[4];
    uint64_t bev = evo[i], bh = hno[i], bv = vno[i];
    bool ok = true;
    for (int c0 = 0; c0 < ne; c0 += kRB) {
        const int e = c0 + threadIdx.x;
        CountSink cs;
        float ax = 0, ay = 0, bx = 0, by = 0;
        if (e < ne) {
            ax = X[g.v + e];
            ay = Y[g.v + e];
            bx = X[g.v + e + 1];
            by = Y[g.v + e + 1];
            walk_edge(g, e, ax, ay, bx, by, cs);
        }
        uint32_t t0, t1, t2;
        const uint32_t o0 = block_scan(cs.nev, sh, &t0);
        const uint32_t o1 = block_scan(cs.nh, sh, &t1);
        const uint32_t o2 = block_scan(cs.nvn, sh, &t2);
        if (e < ne && (cs.nev | cs.nh | cs.nvn)) {
            WriteSink ws;
            ws.g = &g;
            ws.ev = bev + o0;
            ws.h = bh + o1;
            ws.v = bv + o2;
            ws.evkey = evkey;
            ws.evval = evval;
            ws.vnkey = vnkey;
            ws.vnval = vnval;
            ws.hnkey = hnkey;
            ws.hnval = hnval;
            ws.pixev = pixev;
            ws.vlc = vlc;
            ws.hlc = hlc;
            walk_edge(g, e, ax, ay, bx, by, ws);
            ok &= ws.ok;
        }
        bev += t0;
        bh += t1;
        bv += t2;
    }
    if (__syncthreads_or(!ok) && threadIdx.x == 0) atomicOr(&rflag[i], kRejWalk);
}

// ------------------------------------------------------- sequences ----

// Ideal::process_crosses for one pixel's events (in ring order); PASS 0
// counts the sequences, PASS 1 writes them at out.
template <int PASS>
__global__ __launch_bounds__(kRB) void k_rseq(const RPoly *__restrict__ P, const uint32_t *__restrict__ pixpoly,
                                              uint64_t npix, const uint8_t *__restrict__ status,
                                              uint32_t *__restrict__ rflag, const uint64_t *__restrict__ evo,
                                              const uint32_t *__restrict__ pixevo, const uint32_t *__restrict__ ev,
                                              uint32_t *__restrict__ seqc, const uint32_t *__restrict__ seqo,
                                              uint2 *__restrict__ out) {
    for (uint64_t p = (uint64_t)blockIdx.x * kRB + threadIdx.x; p < npix; p += (uint64_t)gridDim.x * kRB) {
        const uint32_t i = pixpoly[p];
        if (rflag[i]) {
            if (PASS == 0) seqc[p] = 0;
            continue;
        }
        const int nv = (int)P[i].nv;
        const uint32_t b = pixevo[p], m = pixevo[p + 1] - b;
        const uint8_t st = status[p];
        bool ok = true;
        uint32_t cnt = 0;
        uint2 *o = PASS ? out + seqo[p] : nullptr;
        auto push = [&](int s, int t) {
            if (t - s + 1 <= 0) ok = false;
            if (PASS) o[cnt] = make_uint2((uint32_t)s, (uint32_t)(t - s + 1));
            cnt++;
        };
        if (m == 0) {
            if (st == kRawBorder) {
                // the whole ring lies in this pixel: the reference records
                // (0, num_vertices), one edge too many; the real edges are kept
                if (evo[i + 1] == evo[i]) push(0, nv - 2);
                else ok = false;  // a BORDER pixel without crossings
            }
        } else {
            const uint32_t *E = ev + b;
            const uint32_t lastv = E[m - 1];
            auto type_at = [&](uint32_t k) -> int { return k < m ? (int)(E[k] & 1u) : (int)!(lastv & 1u); };
            auto edge_at = [&](uint32_t k) -> int { return k < m ? (int)(E[k] >> 1) : (int)(lastv >> 1); };
            const uint32_t mm = m + (m & 1u);
            uint32_t k = 0, last = mm - 1;
            if (type_at(0) == 1) {  // the ring starts inside this pixel
                if (type_at(last) != 0) ok = false;
                push(0, edge_at(0));
                push(edge_at(last), nv - 2);
                k = 1;
                last--;
            }
            for (; k <= last && last != 0xffffffffu; k++) {
                if (type_at(k) != 0) ok = false;
                if (k == last || type_at(k + 1) == 0) {
                    push(edge_at(k), edge_at(k));
                } else {
                    push(edge_at(k), edge_at(k + 1));
                    k++;
                }
            }
        }
        if (PASS == 0) {
            seqc[p] = cnt;
            if (!ok) atomicOr(&rflag[i], kRejSeq);
        }
    }
}

// ------------------------------------------------------- scanline -----

// The relative offset of grid line l as the reference stores it: lines
// without nodes keep 0 (Grid_line.cpp:3-7); line `last` holds the total.
__device__ __forceinline__ uint32_t line_off(const uint32_t *cnt, const uint32_t *off, uint64_t base, int l,
                                            int last, uint32_t total) {
    if (l >= last) return total;
    return cnt[base + l] ? off[base + l] - off[base] : 0u;
}

// Ideal::scanline_reandering: rows >= 1, one lane per row.
__global__ __launch_bounds__(kRB) void k_rscan(const RPoly *__restrict__ P, const uint32_t *__restrict__ rflag,
                                               const uint32_t *__restrict__ hlc, const uint32_t *__restrict__ hlo,
                                               const double *__restrict__ hn, uint8_t *__restrict__ status) {
    const int64_t i = blockIdx.x;
    if (rflag[i]) return;
    const RPoly g = P[i];
    const uint32_t hb = hlo[g.hl];
    const uint32_t total = hlo[g.hl + g.dimy + 2] - hb;
    const double *H = hn + hb;
    for (int y = 1 + threadIdx.x; y < g.dimy; y += kRB) {
        bool inside = false;
        uint32_t a = line_off(hlc, hlo, g.hl, y, g.dimy, total);
        const uint32_t z = line_off(hlc, hlo, g.hl, y + 1, g.dimy, total);
        uint8_t *row = status + g.pix + (uint64_t)y * g.dimx;
        for (int x = 0; x < g.dimx; x++) {
            const uint8_t s = row[x];
            if (s != kRawBorder) {
                row[x] = inside ? kRawIn : 0;
                continue;
            }
            int pass = 0;
            const double right = g.x0 + g.sx * (x + 1);
            while (a < z && H[a] <= right) {
                pass++;
                a++;
            }
            if (pass & 1) inside = !inside;
        }
    }
}

// ---------------------------------------------------- fill classes ----

// Column and row bounds as the reference accumulates them (Xi += step_x),
// and the number of columns / rows its loops visit.
__global__ __launch_bounds__(kRB) void k_rbounds(const RPoly *__restrict__ P, int64_t np,
                                                 uint32_t *__restrict__ rflag, double *__restrict__ colx,
                                                 double *__restrict__ rowy, int *__restrict__ ncols,
                                                 int *__restrict__ nrows) {
    const int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x;
    if (i >= np) return;
    const RPoly g = P[i];
    ncols[i] = nrows[i] = 0;
    if (rflag[i]) return;
    bool over = false;
    {
        double Xi = g.x0, Xi1 = Xi + g.sx;
        const double kx = Xi + g.dimx * g.sx;
        int c = 0;
        while (Xi1 < kx + 1e-9) {
            if (c > g.dimx) {
                over = true;
                break;
            }
            colx[g.vl + c] = Xi;
            Xi += g.sx;
            Xi1 = Xi + g.sx;
            c++;
        }
        colx[g.vl + c] = Xi;
        ncols[i] = c;
    }
    {
        double Yi = g.y0, Yi1 = Yi + g.sy;
        const double ky = Yi + g.dimy * g.sy;
        int r = 0;
        while (Yi1 < ky + 1e-9) {
            if (r > g.dimy) {
                over = true;
                break;
            }
            rowy[g.hl + r] = Yi;
            Yi += g.sy;
            Yi1 = Yi + g.sy;
            r++;
        }
        rowy[g.hl + r] = Yi;
        nrows[i] = r;
    }
    if (over) atomicOr(&rflag[i], kRejUnsupported);
}

// The host's bucket of cells an edge can reach along one axis (its margin
// covers every rounding of the clip tests); [c0, c1] of the cells.
struct Bucket {
    double lo, inv, m;
    int ncell;
    __device__ Bucket(double lo_, double step, int ncell_, double vmax) : lo(lo_), ncell(ncell_) {
        inv = 1.0 / step;
        double amax = fmax(fabs(lo), fabs(lo + (ncell + 1) * step));
        amax = (amax < vmax) ? vmax : amax;
        m = amax * (0x1p-22 + (ncell + 8) * 0x1p-50) + step * 0x1p-20;
    }
    __device__ __forceinline__ bool reaches(float a, float b, int c) const {
        const double mn = smin(a, b), mx = smax(a, b);
        const int c0 = max(cvt_x86(floor((mn - m - lo) * inv)), 0);
        const int c1 = min(cvt_x86(floor((mx + m - lo) * inv)), ncell - 1);
        return c0 <= c && c <= c1;
    }
};

// Points edge e adds to column strip [Xi, Xi1] (Ideal::calculate_fullness,
// its x pass): vertex A (edge 0 only: A of a later edge repeats the
// previous edge's B and is always a duplicate), the crossings with the two
// column bounds, vertex B.
__device__ __forceinline__ int strip_points(float Ax, float Ay, float Bx, float By, bool first, double Xi, double Xi1,
                                            float2 *out) {
    double yi, yi1;
    if (Ax == Bx) {
        yi = smin(Ay, By);
        yi1 = yi;
    } else {
        const float den = Bx - Ax;
        const double slope = ((double)By - Ay) / den;
        yi = slope * (Xi - Bx) + By;
        yi1 = slope * (Xi1 - Bx) + By;
    }
    const float2 q0 = make_float2((float)Xi, (float)yi), q1 = make_float2((float)Xi1, (float)yi1);
    const float mny = smin(Ay, By), mxy = smax(Ay, By);
    const float mnx = smin(Ax, Bx), mxx = smax(Ax, Bx);
    const bool in0 = within(mny, mxy, q0.y) && within(mnx, mxx, q0.x);
    const bool in1 = within(mny, mxy, q1.y) && within(mnx, mxx, q1.x);
    // constant indexes only, so that out[] stays in registers
    int n = 0;
    auto app = [&](float2 p) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (n == j) out[j] = p;
        n++;
    };
    if (first && within(Xi, Xi1, Ax)) app(make_float2(Ax, Ay));
    if (Ax < Bx) {
        if (in0) app(q0);
        if (in1) app(q1);
    } else {
        if (in1) app(q1);
        if (in0) app(q0);
    }
    if (within(Xi, Xi1, Bx)) app(make_float2(Bx, By));
    return n;
}

__device__ __forceinline__ unsigned long long pt_key(float2 p) {
    return ((unsigned long long)__float_as_uint(p.x) << 32) | __float_as_uint(p.y);
}

// Insert key with first-occurrence index idx; returns its slot.
__device__ __forceinline__ uint32_t hash_put(unsigned long long *keys, uint32_t *ix, uint32_t mask,
                                             unsigned long long key, uint32_t idx) {
    uint32_t s = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
    for (;;) {
        unsigned long long k = keys[s];
        if (k == 0) {
            k = atomicCAS(&keys[s], 0ull, key);
            if (k == 0) k = key;
        }
        if (k == key) {
            atomicMin(&ix[s], idx);
            return s;
        }
        s = (s + 1) & mask;
    }
}

// Candidate counts of every column strip (PASS 0, before dedup) or the
// strips themselves (PASS 1): kept points in order, the first repeated at
// the end, their count and largest |y|.
template <int PASS>
__global__ __launch_bounds__(kRB) void k_rstrip(const RPoly *__restrict__ P, const float *__restrict__ X,
                                                const float *__restrict__ Y, const float *__restrict__ absx,
                                                uint32_t *__restrict__ rflag, const int *__restrict__ ncols,
                                                const int *__restrict__ nrows, const double *__restrict__ colx,
                                                uint32_t *__restrict__ stripc, const uint64_t *__restrict__ stripo,
                                                const uint64_t *__restrict__ hasho, float2 *__restrict__ pool,
                                                unsigned long long *__restrict__ hkeys, uint32_t *__restrict__ hidx,
                                                uint32_t *__restrict__ stripn, float *__restrict__ stripmy) {
    const int64_t i = blockIdx.x;
    if (rflag[i]) return;
    const RPoly g = P[i];
    const int ncol = ncols[i], nrow = nrows[i];
    const Bucket bk(g.x0, g.sx, g.dimx + 2, (double)absx[i]);
    const float *Xp = X + g.v, *Yp = Y + g.v;
    const int ne = (int)g.nv - 1;
    __shared__ uint32_t sh[4];
    __shared__ unsigned long long s_keys[PASS ? kLdsSlots : 1];
    __shared__ uint32_t s_idx[PASS ? kLdsSlots : 1];
    __shared__ float2 s_cand[PASS ? kChunkCands : 1];
    __shared__ uint8_t s_keep[PASS ? kChunkCands : 1], s_near[PASS ? kChunkCands : 1];
    __shared__ float s_my[kRB / 64];
    bool fail = false;
    for (int c = 0; c < ncol; c++) {
        const double Xi = colx[g.vl + c], Xi1 = colx[g.vl + c + 1];
        if (PASS == 0) {
            uint32_t cnt = 0;
            for (int e = threadIdx.x; e < ne; e += kRB) {
                const float ax = Xp[e], bx = Xp[e + 1];
                if (!bk.reaches(ax, bx, c)) continue;
                float2 pts[4];
                cnt += strip_points(ax, Yp[e], bx, Yp[e + 1], e == 0, Xi, Xi1, pts);
            }
            uint32_t tot;
            block_scan(cnt, sh, &tot);
            if (threadIdx.x == 0) stripc[g.vl + c] = tot;
            continue;
        }
        const uint32_t cnt = stripc[g.vl + c];
        if (cnt == 0) {
            if (threadIdx.x == 0) stripn[g.vl + c] = 0;
            continue;
        }
        float2 *pts = pool + stripo[g.vl + c];
        const bool lds = cnt <= kLdsCands;
        uint32_t slots = kLdsSlots;
        if (!lds) {
            slots = 1;
            while (slots < 2 * cnt) slots <<= 1;
        }
        unsigned long long *gkeys = hkeys + (lds ? 0 : hasho[g.vl + c]);
        uint32_t *gix = hidx + (lds ? 0 : hasho[g.vl + c]);
        const uint32_t mask = slots - 1;
        if (lds) {
            for (uint32_t s = threadIdx.x; s < slots; s += kRB) {
                s_keys[s] = 0;
                s_idx[s] = 0xffffffffu;
            }
        } else {
            for (uint32_t s = threadIdx.x; s < slots; s += kRB) {
                gkeys[s] = 0;
                gix[s] = 0xffffffffu;
            }
        }
        __syncthreads();
        uint32_t base = 0, kept = 0;
        float my = 0.f;
        for (int c0 = 0; c0 < ne; c0 += kRB) {
            const int e = c0 + threadIdx.x;
            float2 loc[4];
            uint32_t n = 0;
            if (e < ne) {
                const float ax = Xp[e], bx = Xp[e + 1];
                if (bk.reaches(ax, bx, c)) n = strip_points(ax, Yp[e], bx, Yp[e + 1], e == 0, Xi, Xi1, loc);
            }
            uint32_t tot;
            const uint32_t off = block_scan(n, sh, &tot);
            if (tot == 0) continue;
            uint32_t slot[4];
            bool anynear = false;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                if (j >= n) break;
                s_cand[off + j] = loc[j];
                const bool nr = near_axis(loc[j]);
                s_near[off + j] = nr;
                anynear |= nr;
                if (!nr)
                    slot[j] = lds ? hash_put(s_keys, s_idx, mask, pt_key(loc[j]), base + off + j)
                                  : hash_put(gkeys, gix, mask, pt_key(loc[j]), base + off + j);
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                if (j >= n) break;
                s_keep[off + j] =
                    s_near[off + j] ? 0 : ((lds ? s_idx[slot[j]] : gix[slot[j]]) == base + off + j);
            }
            if (__syncthreads_or(anynear)) {
                // points near an axis, in order, against every point kept before them
                for (uint32_t j = 0; j < tot; j++) {
                    if (!s_near[j]) continue;
                    const float2 p = s_cand[j];
                    bool found = false;
                    for (uint32_t k = threadIdx.x; k < kept; k += kRB) found |= same_pt(pts[k], p);
                    for (uint32_t k = threadIdx.x; k < j; k += kRB) found |= s_keep[k] && same_pt(s_cand[k], p);
                    found = __syncthreads_or(found);
                    if (threadIdx.x == 0) s_keep[j] = !found;
                    __syncthreads();
                }
            }
            uint32_t nk = 0;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++)
                if (j < n) nk += s_keep[off + j];
            uint32_t ktot;
            uint32_t ko = block_scan(nk, sh, &ktot);
#pragma unroll
            for (uint32_t j = 0; j < 4; j++)
                if (j < n && s_keep[off + j]) {
                    pts[kept + ko++] = loc[j];
                    my = smax(my, fabsf(loc[j].y));
                }
            base += tot;
            kept += ktot;
        }
        // the largest |y| of the strip (the row buckets' margin scale)
        const float wm = wave_max(my);
        if ((threadIdx.x & 63) == 0) s_my[threadIdx.x >> 6] = wm;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = s_my[0];
            for (int w = 1; w < kRB / 64; w++) m = smax(m, s_my[w]);
            if (kept > 0) pts[kept] = pts[0];
            stripn[g.vl + c] = kept ? kept + 1 : 0;
            stripmy[g.vl + c] = m;
            // a non-empty strip outside the grid: the reference asserts
            if (kept > 0 && (c >= g.dimx || nrow > g.dimy)) fail = true;
        }
        __syncthreads();
    }
    if (PASS == 1 && threadIdx.x == 0 && fail) atomicOr(&rflag[i], kRejFill);
}

// Ideal::calculate_fullness, y pass and encodePixelArea (Ideal.cpp:64-88),
// one lane per pixel of a visited cell.
__global__ __launch_bounds__(kRB) void k_rcells(const RPoly *__restrict__ P, const uint32_t *__restrict__ pixpoly,
                                                uint64_t npix, const uint32_t *__restrict__ rflag,
                                                const int *__restrict__ ncols, const int *__restrict__ nrows,
                                                const double *__restrict__ rowy, const uint64_t *__restrict__ stripo,
                                                const uint32_t *__restrict__ stripn,
                                                const float *__restrict__ stripmy, const float2 *__restrict__ pool,
                                                uint8_t *__restrict__ status, double *__restrict__ areas) {
    for (uint64_t p = (uint64_t)blockIdx.x * kRB + threadIdx.x; p < npix; p += (uint64_t)gridDim.x * kRB) {
        const uint32_t i = pixpoly[p];
        if (rflag[i]) continue;
        const RPoly g = P[i];
        const uint64_t local = p - g.pix;
        const int col = (int)(local % (uint64_t)g.dimx), row = (int)(local / (uint64_t)g.dimx);
        if (col >= ncols[i] || row >= nrows[i]) continue;
        const uint32_t n = stripn[g.vl + col];
        if (n == 0) continue;
        const uint8_t st = status[p];
        const double pix_area = g.sx * g.sy;
        if (st == kRawIn) {
            status[p] = kCC - 1;
            areas[p] = pix_area;
            continue;
        }
        if (st != kRawBorder) {
            status[p] = 0;
            areas[p] = 0.0;
            continue;
        }
        const double Yi = rowy[g.hl + row], Yi1 = rowy[g.hl + row + 1];
        const Bucket bk(g.y0, g.sy, g.dimy + 2, (double)stripmy[g.vl + col]);
        const float2 *V = pool + stripo[g.vl + col];
        // the cell's points, summed into the shoelace as they come
        double area = 0.0;
        uint32_t cnt = 0;
        float2 first = make_float2(0.f, 0.f), prv = first;
        auto push = [&](float2 q) {
            if (cnt == 0) first = q;
            else area += ((double)prv.x * q.y - (double)q.x * prv.y);
            prv = q;
            cnt++;
        };
        float2 A = V[0];
        for (uint32_t k = 0; k + 1 < n; k++) {
            const float2 B = V[k + 1];
            if (bk.reaches(A.y, B.y, row)) {
                bool skip = false, ok0 = true, ok1 = true;
                double xi = 0, xi1 = 0;
                if (A.x == B.x) {
                    xi = A.x;
                    xi1 = A.x;
                } else {
                    const float den = B.x - A.x;
                    const double slope = ((double)B.y - A.y) / den;
                    if (A.y == B.y) {
                        if (A.y == Yi) {
                            xi = A.x;
                            ok1 = false;
                        } else if (A.y == Yi1) {
                            ok0 = false;
                            xi1 = A.x;
                        } else {
                            skip = true;
                        }
                    } else {
                        xi = ((Yi - B.y) / slope) + B.x;
                        xi1 = ((Yi1 - B.y) / slope) + B.x;
                    }
                }
                if (within(Yi, Yi1, A.y)) push(A);
                if (!skip) {
                    const float2 q0 = make_float2((float)xi, (float)Yi), q1 = make_float2((float)xi1, (float)Yi1);
                    const float mny = smin(A.y, B.y), mxy = smax(A.y, B.y);
                    const float mnx = smin(A.x, B.x), mxx = smax(A.x, B.x);
                    const bool in0 = ok0 && within(mnx, mxx, q0.x) && within(mny, mxy, q0.y);
                    const bool in1 = ok1 && within(mnx, mxx, q1.x) && within(mny, mxy, q1.y);
                    if (A.y < B.y) {
                        if (in0) push(q0);
                        if (in1) push(q1);
                    } else {
                        if (in1) push(q1);
                        if (in0) push(q0);
                    }
                }
            }
            A = B;
        }
        if (cnt > 0 && !same_pt(first, prv)) push(first);
        area = cnt >= 3 ? fabs(area) / 2.0 : 0.0;
        int cls;
        const double ratio = area / pix_area;
        if (fabs(ratio - 1.0) < 1e-9) cls = kCC - 1;
        else if (fabs(ratio) < 1e-9) cls = 0;
        else {
            cls = cvt_x86(ceil(ratio * (kCC - 2)));
            if (cls >= kCC) cls = kCC - 1;
            cls = (uint8_t)cls;
        }
        status[p] = (uint8_t)max(1, min(cls, kCC - 2));
        areas[p] = area;
    }
}

// --------------------------------------------------------- packing ----

__global__ __launch_bounds__(kRB) void k_rcounts(const RPoly *__restrict__ P, int64_t np,
                                                 uint32_t *__restrict__ rflag, const uint32_t *__restrict__ seqo,
                                                 const uint32_t *__restrict__ vlo, uint64_t *__restrict__ nseq,
                                                 uint64_t *__restrict__ nvn) {
    const int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x;
    if (i > np) return;
    if (i == np) {  // the scans' last entry
        nseq[i] = nvn[i] = 0;
        return;
    }
    const RPoly g = P[i];
    const uint64_t npx = (uint64_t)g.dimx * g.dimy;
    const uint32_t s = seqo[g.pix + npx] - seqo[g.pix];
    if (s >= 65536) atomicOr(&rflag[i], kRejSeq);  // Ideal.cpp:32
    const bool rej = rflag[i] != 0;
    nseq[i] = rej ? 0 : s;
    nvn[i] = rej ? 0 : vlo[g.vl + g.dimx + 2] - vlo[g.vl];
}

// Final arrays of one polygon: header fields, sequences, pixel and grid-line
// offsets (relative), nodes; a rejected polygon is an empty raster.
__global__ __launch_bounds__(kRB) void k_rpack(const RPoly *__restrict__ P, const uint32_t *__restrict__ rflag,
                                               const uint64_t *__restrict__ seqf, const uint64_t *__restrict__ vnf,
                                               const uint32_t *__restrict__ seqo, const uint2 *__restrict__ seqtmp,
                                               const uint32_t *__restrict__ vlc, const uint32_t *__restrict__ vlo,
                                               const double *__restrict__ vn, farm_raster_hdr *__restrict__ hdr,
                                               uint8_t *__restrict__ status, double *__restrict__ areas,
                                               uint32_t *__restrict__ pixoff, uint2 *__restrict__ seqs,
                                               uint32_t *__restrict__ voff, double *__restrict__ vnodes) {
    const int64_t i = blockIdx.x;
    const RPoly g = P[i];
    const bool rej = rflag[i] != 0;
    const uint64_t npx = (uint64_t)g.dimx * g.dimy;
    const uint64_t po = hdr[i].pixoff_start;
    const uint32_t ns = (uint32_t)(seqf[i + 1] - seqf[i]), nn = (uint32_t)(vnf[i + 1] - vnf[i]);
    if (threadIdx.x == 0) {
        hdr[i].seq_start = seqf[i];
        hdr[i].nseqs = ns;
        hdr[i].vnode_start = vnf[i];
        hdr[i].flags = rej ? FARM_HDR_REJECTED : 0;
    }
    const int nl = g.dimx + 2;
    if (rej) {
        for (uint64_t p = threadIdx.x; p < npx; p += kRB) {
            status[g.pix + p] = 0;
            areas[g.pix + p] = 0.0;
        }
        for (uint64_t p = threadIdx.x; p <= npx; p += kRB) pixoff[po + p] = 0;
        for (int l = threadIdx.x; l < nl; l += kRB) voff[g.vl + l] = 0;
        return;
    }
    const uint32_t s0 = seqo[g.pix];
    for (uint32_t k = threadIdx.x; k < ns; k += kRB) seqs[seqf[i] + k] = seqtmp[s0 + k];
    for (uint64_t p = threadIdx.x; p <= npx; p += kRB) pixoff[po + p] = seqo[g.pix + p] - s0;
    const uint32_t vb = vlo[g.vl];
    for (int l = threadIdx.x; l < nl; l += kRB) voff[g.vl + l] = line_off(vlc, vlo, g.vl, l, g.dimx, nn);
    for (uint32_t k = threadIdx.x; k < nn; k += kRB) vnodes[vnf[i] + k] = vn[vb + k];
}

// Ideal::merge_status for every coarse pixel of every layer.
__global__ __launch_bounds__(kRB) void k_rmerge(const farm_raster_hdr *__restrict__ hdr,
                                                const farm_layer *__restrict__ layers,
                                                const uint32_t *__restrict__ lpoly, uint64_t nlayers, uint64_t nls,
                                                const uint8_t *__restrict__ status,
                                                const double *__restrict__ areas, uint8_t *__restrict__ lstatus) {
    for (uint64_t q = (uint64_t)blockIdx.x * kRB + threadIdx.x; q < nls; q += (uint64_t)gridDim.x * kRB) {
        uint64_t lo = 0, hi = nlayers - 1;  // the last layer starting at or before q
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (layers[mid].status_start <= q) lo = mid;
            else hi = mid - 1;
        }
        const farm_layer L = layers[lo];
        const farm_raster_hdr &f = hdr[lpoly[lo]];
        const uint64_t local = q - L.status_start;
        const int x = (int)(local % (uint64_t)L.dimx), y = (int)(local / (uint64_t)L.dimx);
        const double s0 = L.step, s1 = L.step;
        const double lx = L.low[0] + x * s0, ly = L.low[1] + y * s1;
        const double hx = L.low[0] + (x + 1) * s0, hy = L.low[1] + (y + 1) * s1;
        auto grid_off = [](double v, double lo_, double step, int dim) {
            return min(max(dbl2int_x86((v - lo_) / step), 0), dim - 1);
        };
        const int x0 = grid_off(lx, f.mbr[0], f.step[0], f.dimx);
        const int y0 = grid_off(ly, f.mbr[1], f.step[1], f.dimy);
        int x1 = grid_off(hx, f.mbr[0], f.step[0], f.dimx) - 1;
        int y1 = grid_off(hy, f.mbr[1], f.step[1], f.dimy) - 1;
        if (y1 < y0) y1 = y0;
        if (x1 < x0) x1 = x0;
        double area = 0.0;
        bool border = false;
        const uint8_t *st = status + f.pix_start;
        const double *ar = areas + f.pix_start;
        for (int a = x0; a <= x1; a++)
            for (int b = y0; b <= y1; b++) {
                const uint8_t c = st[(uint64_t)b * f.dimx + a];
                if (c != 0 && c != kCC - 1) border = true;
                area += ar[(uint64_t)b * f.dimx + a];
            }
        uint8_t cls;
        if (border) {
            const double ratio = area / (s0 * s1);
            if (fabs(ratio - 1.0) < 1e-9) cls = kCC - 1;
            else if (fabs(ratio) < 1e-9) cls = 0;
            else {
                int idx = cvt_x86(ceil(ratio * (kCC - 2)));
                if (idx >= kCC) idx = kCC - 1;
                cls = (uint8_t)idx;
            }
        } else {
            const uint8_t c = st[(uint64_t)y0 * f.dimx + x0];
            cls = (c == kCC - 1) ? kCC - 1 : 0;
        }
        lstatus[q] = cls;
    }
}

// BORDER pixels per polygon (the within kernel's LDS sizing).
__global__ __launch_bounds__(kRB) void k_rborder_max(const RPoly *__restrict__ P, const uint8_t *__restrict__ status,
                                                     unsigned long long *__restrict__ out) {
    const RPoly g = P[blockIdx.x];
    const uint64_t npx = (uint64_t)g.dimx * g.dimy;
    uint32_t nb = 0;
    for (uint64_t p = threadIdx.x; p < npx; p += kRB) {
        const uint8_t c = status[g.pix + p];
        nb += (c != 0 && c != kCC - 1);
    }
    __shared__ uint32_t sh[4];
    uint32_t tot;
    block_scan(nb, sh, &tot);
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)tot);
}

// -w b: fill classes at cc categories from the clipped areas
// (encodePixelArea, Ideal.cpp:64-88, with cc in place of 16, and the BORDER
// clamp of calculate_fullness, Ideal.cpp:905-912); one workgroup per polygon.
__global__ __launch_bounds__(kRB) void k_rfill(const farm_raster_hdr *__restrict__ hdr,
                                               const uint8_t *__restrict__ status,
                                               const double *__restrict__ areas, int cc,
                                               uint8_t *__restrict__ fill) {
    const farm_raster_hdr h = hdr[blockIdx.x];
    const double pix_area = h.step[0] * h.step[1];
    const uint64_t npx = (uint64_t)h.dimx * h.dimy;
    for (uint64_t p = threadIdx.x; p < npx; p += kRB) {
        const uint8_t st = status[h.pix_start + p];
        uint8_t f = 0;
        if (st == kCC - 1) {
            f = (uint8_t)(cc - 1);
        } else if (st != 0) {
            const double ratio = areas[h.pix_start + p] / pix_area;
            int c;
            if (fabs(ratio - 1.0) < 1e-9) c = cc - 1;
            else if (fabs(ratio) < 1e-9) c = 0;
            else {
                c = cvt_x86(ceil(ratio * (cc - 2)));
                if (c >= cc) c = cc - 1;
            }
            f = (uint8_t)max(1, min(c, cc - 2));
        }
        fill[h.pix_start + p] = f;
    }
}

// Device temporaries of one build, freed when it returns.
struct Tmp {
    std::vector<void *> bufs;
    ~Tmp() {
        for (void *p : bufs) (void)hipFree(p);
    }
    template <class T>
    int get(farm_handle h, T **out, uint64_t n) {
        void *p = nullptr;
        const size_t bytes = std::max<size_t>(16, sizeof(T) * n);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            return fail(h, FARM_ENOMEM, "raster builder: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        bufs.push_back(p);
        *out = (T *)p;
        return FARM_OK;
    }
};

template <class T>
int scan_excl(farm_handle h, const T *in, T *out, uint64_t n) {
    if (n >= (1ull << 31)) return fail(h, FARM_ECAPACITY, "raster builder: scan of %llu entries", (unsigned long long)n);
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, h->stream));
    int rc = ensure(h, h->scan_tmp, tmp);
    if (rc) return rc;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp.p, tmp, in, out, (int)n, h->stream));
    return FARM_OK;
}

template <class K, class V>
int sort_pairs(farm_handle h, const K *kin, K *kout, const V *vin, V *vout, uint64_t n, int end_bit) {
    if (n == 0) return FARM_OK;
    if (n >= (1ull << 31)) return fail(h, FARM_ECAPACITY, "raster builder: sort of %llu entries", (unsigned long long)n);
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, (int)n, 0, end_bit, h->stream));
    int rc = ensure(h, h->scan_tmp, tmp);
    if (rc) return rc;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(h->scan_tmp.p, tmp, kin, kout, vin, vout, (int)n, 0, end_bit,
                                                 h->stream));
    return FARM_OK;
}

int bits_for(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
}

template <class T>
int read1(farm_handle h, const T *dev, T *host) {
    HIPCHK(h, hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

unsigned grid_for(uint64_t n) { return (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (n + kRB - 1) / kRB), 65536); }

}  // namespace

}  // namespace farm

using namespace farm;

namespace farm {
namespace {

// Everything after the vertex upload: x / y of the handle hold the closed
// rings, ring i at [voff[i], voff[i + 1]).
int rasterize_uploaded(farm_handle h, int64_t np, const std::vector<uint64_t> &voff, uint64_t nvt, int vpr,
                       int nlow, farm_raster_info *info) {
    const bool hier = nlow > 0;
    int rc;
    Tmp T;
    uint64_t *d_voff;
    double *d_mbb;
    float *d_absx;
    if ((rc = T.get(h, &d_voff, np + 1)) || (rc = T.get(h, &d_mbb, 4 * np)) || (rc = T.get(h, &d_absx, np)))
        return rc;
    HIPCHK(h, hipMemcpyAsync(d_voff, voff.data(), 8 * (np + 1), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    std::vector<double> mbb(4 * np);
    if (np) {
        hipLaunchKernelGGL(k_rmbb, dim3((unsigned)np), dim3(kRB), 0, h->stream, P<float>(h->x), P<float>(h->y),
                           (const uint64_t *)d_voff, d_mbb, d_absx);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(mbb.data(), d_mbb, 32 * np, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));

    // grids and array offsets on the host (the reference's pow / log / round)
    std::vector<farm_raster_hdr> hdr(np);
    std::vector<RPoly> rp(np);
    std::vector<uint32_t> flag0(np, 0);
    std::vector<farm_layer> layers;
    std::vector<uint32_t> lpoly;
    uint64_t npix = 0, npo = 0, nvl = 0, nhl = 0, nls = 0, max_layer_pix = 0;
    std::vector<RasterGrid> coarse;
    for (int64_t i = 0; i < np; i++) {
        const int nv = (int)(voff[i + 1] - voff[i]);
        RasterGrid g;
        bool ok = true;
        if (hier) ok = hier_grids(&mbb[4 * i], nv, vpr, nlow, g, coarse);
        else g = size_grid(&mbb[4 * i], nv / vpr);
        const uint64_t n = (uint64_t)g.dimx * (uint64_t)g.dimy;
        if (g.dimx <= 0 || g.dimy <= 0 || n >= (1ull << 31))
            return fail(h, FARM_EINVAL, "polygon %lld: raster grid %d x %d", (long long)i, g.dimx, g.dimy);
        farm_raster_hdr &r = hdr[i];
        memset(&r, 0, sizeof r);
        r.mbr[0] = g.low[0];
        r.mbr[1] = g.low[1];
        r.mbr[2] = g.high[0];
        r.mbr[3] = g.high[1];
        r.step[0] = g.step[0];
        r.step[1] = g.step[1];
        r.dimx = g.dimx;
        r.dimy = g.dimy;
        r.pix_start = npix;
        r.pixoff_start = npo;
        r.vert_start = voff[i];
        r.nverts = (uint32_t)nv;
        r.vline_start = nvl;
        RPoly &q = rp[i];
        q.x0 = g.low[0];
        q.y0 = g.low[1];
        q.sx = g.step[0];
        q.sy = g.step[1];
        q.dimx = g.dimx;
        q.dimy = g.dimy;
        q.pix = npix;
        q.vl = nvl;
        q.hl = nhl;
        q.v = voff[i];
        q.nv = (uint32_t)nv;
        q.pad = 0;
        if (!ok) flag0[i] = kRejAlign;
        max_layer_pix = std::max(max_layer_pix, n);
        if (hier) {
            r.layer_start = layers.size();
            r.nlayers = ok ? (int32_t)coarse.size() + 1 : 1;
            if (ok)
                for (const RasterGrid &c : coarse) {
                    farm_layer l;
                    l.low[0] = c.low[0];
                    l.low[1] = c.low[1];
                    l.step = c.step[0];
                    l.dimx = c.dimx;
                    l.dimy = c.dimy;
                    l.status_start = nls;
                    layers.push_back(l);
                    lpoly.push_back((uint32_t)i);
                    const uint64_t cn = (uint64_t)c.dimx * c.dimy;
                    nls += cn;
                    max_layer_pix = std::max(max_layer_pix, cn);
                }
        }
        npix += n;
        npo += n + 1;
        nvl += (uint64_t)g.dimx + 2;
        nhl += (uint64_t)g.dimy + 2;
    }
    if (npix + 1 >= (1ull << 31) || nvl + 1 >= (1ull << 31) || nhl + 1 >= (1ull << 31))
        return fail(h, FARM_ECAPACITY, "raster builder: %llu pixels / %llu + %llu grid lines exceed 2^31",
                    (unsigned long long)npix, (unsigned long long)nvl, (unsigned long long)nhl);

    RPoly *d_rp;
    uint32_t *d_flag, *d_pixpoly;
    if ((rc = T.get(h, &d_rp, np)) || (rc = T.get(h, &d_flag, np)) || (rc = T.get(h, &d_pixpoly, npix))) return rc;
    HIPCHK(h, hipMemcpyAsync(d_rp, rp.data(), sizeof(RPoly) * np, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(d_flag, flag0.data(), 4 * np, hipMemcpyHostToDevice, h->stream));
    if ((rc = upload(h, h->hdr, hdr.data(), sizeof(farm_raster_hdr) * np))) return rc;
    if ((rc = ensure(h, h->status, npix))) return rc;
    if ((rc = ensure(h, h->areas, 8 * npix))) return rc;
    HIPCHK(h, hipMemsetAsync(h->status.p, 0, npix, h->stream));
    HIPCHK(h, hipMemsetAsync(h->areas.p, 0, 8 * npix, h->stream));
    if (np) {
        hipLaunchKernelGGL(k_rinit, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp, d_pixpoly);
        HIPCHK(h, hipGetLastError());
    }

    // ---- edge walk: count, scan, write ----
    uint64_t *d_evc, *d_hnc, *d_vnc, *d_evo, *d_hno, *d_vno;
    if ((rc = T.get(h, &d_evc, np + 1)) || (rc = T.get(h, &d_hnc, np + 1)) || (rc = T.get(h, &d_vnc, np + 1)) ||
        (rc = T.get(h, &d_evo, np + 1)) || (rc = T.get(h, &d_hno, np + 1)) || (rc = T.get(h, &d_vno, np + 1)))
        return rc;
    HIPCHK(h, hipMemsetAsync(d_evc + np, 0, 8, h->stream));
    HIPCHK(h, hipMemsetAsync(d_hnc + np, 0, 8, h->stream));
    HIPCHK(h, hipMemsetAsync(d_vnc + np, 0, 8, h->stream));
    if (np) {
        hipLaunchKernelGGL(k_rwalk_count, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           P<float>(h->x), P<float>(h->y), P<uint8_t>(h->status), d_flag, d_evc, d_hnc, d_vnc);
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = scan_excl(h, (const uint64_t *)d_evc, d_evo, np + 1)) ||
        (rc = scan_excl(h, (const uint64_t *)d_hnc, d_hno, np + 1)) ||
        (rc = scan_excl(h, (const uint64_t *)d_vnc, d_vno, np + 1)))
        return rc;
    uint64_t nev = 0, nhn = 0, nvn = 0;
    if ((rc = read1(h, d_evo + np, &nev)) || (rc = read1(h, d_hno + np, &nhn)) || (rc = read1(h, d_vno + np, &nvn)))
        return rc;
    if (nev >= (1ull << 31) || nhn >= (1ull << 31) || nvn >= (1ull << 31))
        return fail(h, FARM_ECAPACITY, "raster builder: %llu crossings exceed 2^31; rasterize fewer polygons",
                    (unsigned long long)nev);
    uint32_t *evk, *evv, *evk2, *evv2, *vnk, *vnk2, *hnk, *hnk2, *pixev, *pixevo, *vlc, *vlo, *hlc, *hlo;
    double *vnv, *vnv2, *hnv, *hnv2;
    if ((rc = T.get(h, &evk, nev)) || (rc = T.get(h, &evv, nev)) || (rc = T.get(h, &evk2, nev)) ||
        (rc = T.get(h, &evv2, nev)) || (rc = T.get(h, &vnk, nvn)) || (rc = T.get(h, &vnk2, nvn)) ||
        (rc = T.get(h, &vnv, nvn)) || (rc = T.get(h, &vnv2, nvn)) || (rc = T.get(h, &hnk, nhn)) ||
        (rc = T.get(h, &hnk2, nhn)) || (rc = T.get(h, &hnv, nhn)) || (rc = T.get(h, &hnv2, nhn)) ||
        (rc = T.get(h, &pixev, npix + 1)) || (rc = T.get(h, &pixevo, npix + 1)) || (rc = T.get(h, &vlc, nvl + 1)) ||
        (rc = T.get(h, &vlo, nvl + 1)) || (rc = T.get(h, &hlc, nhl + 1)) || (rc = T.get(h, &hlo, nhl + 1)))
        return rc;
    HIPCHK(h, hipMemsetAsync(pixev, 0, 4 * (npix + 1), h->stream));
    HIPCHK(h, hipMemsetAsync(vlc, 0, 4 * (nvl + 1), h->stream));
    HIPCHK(h, hipMemsetAsync(hlc, 0, 4 * (nhl + 1), h->stream));
    if (np) {
        hipLaunchKernelGGL(k_rwalk_write, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           P<float>(h->x), P<float>(h->y), d_flag, (const uint64_t *)d_evo, (const uint64_t *)d_hno,
                           (const uint64_t *)d_vno, evk, evv, vnk, vnv, hnk, hnv, pixev, vlc, hlc);
        HIPCHK(h, hipGetLastError());
    }
    // events in (pixel, ring) order; nodes in (line, value) order
    if ((rc = sort_pairs(h, (const uint32_t *)evk, evk2, (const uint32_t *)evv, evv2, nev, bits_for(npix + 1))))
        return rc;
    if ((rc = sort_pairs(h, (const double *)vnv, vnv2, (const uint32_t *)vnk, vnk2, nvn, 64)) ||
        (rc = sort_pairs(h, (const uint32_t *)vnk2, vnk, (const double *)vnv2, vnv, nvn, bits_for(nvl + 1))))
        return rc;
    if ((rc = sort_pairs(h, (const double *)hnv, hnv2, (const uint32_t *)hnk, hnk2, nhn, 64)) ||
        (rc = sort_pairs(h, (const uint32_t *)hnk2, hnk, (const double *)hnv2, hnv, nhn, bits_for(nhl + 1))))
        return rc;
    if ((rc = scan_excl(h, (const uint32_t *)pixev, pixevo, npix + 1)) ||
        (rc = scan_excl(h, (const uint32_t *)vlc, vlo, nvl + 1)) ||
        (rc = scan_excl(h, (const uint32_t *)hlc, hlo, nhl + 1)))
        return rc;

    // ---- edge sequences ----
    uint32_t *seqc, *seqo;
    if ((rc = T.get(h, &seqc, npix + 1)) || (rc = T.get(h, &seqo, npix + 1))) return rc;
    HIPCHK(h, hipMemsetAsync(seqc + npix, 0, 4, h->stream));
    if (npix) {
        hipLaunchKernelGGL(k_rseq<0>, dim3(grid_for(npix)), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           (const uint32_t *)d_pixpoly, npix, P<uint8_t>(h->status), d_flag, (const uint64_t *)d_evo,
                           (const uint32_t *)pixevo, (const uint32_t *)evv2, seqc, (const uint32_t *)nullptr,
                           (uint2 *)nullptr);
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = scan_excl(h, (const uint32_t *)seqc, seqo, npix + 1))) return rc;
    uint32_t nseq_tmp = 0;
    if ((rc = read1(h, seqo + npix, &nseq_tmp))) return rc;
    uint2 *seqtmp;
    if ((rc = T.get(h, &seqtmp, nseq_tmp))) return rc;
    if (npix) {
        hipLaunchKernelGGL(k_rseq<1>, dim3(grid_for(npix)), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           (const uint32_t *)d_pixpoly, npix, P<uint8_t>(h->status), d_flag, (const uint64_t *)d_evo,
                           (const uint32_t *)pixevo, (const uint32_t *)evv2, seqc, (const uint32_t *)seqo, seqtmp);
        HIPCHK(h, hipGetLastError());
    }

    // ---- IN / OUT ----
    if (np) {
        hipLaunchKernelGGL(k_rscan, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           (const uint32_t *)d_flag, (const uint32_t *)hlc, (const uint32_t *)hlo,
                           (const double *)hnv, P<uint8_t>(h->status));
        HIPCHK(h, hipGetLastError());
    }

    // ---- fill classes ----
    double *colx, *rowy;
    int *ncols, *nrows;
    uint32_t *stripc, *stripn;
    uint64_t *psz, *stripo, *hsz, *hasho;
    float *stripmy;
    if ((rc = T.get(h, &colx, nvl)) || (rc = T.get(h, &rowy, nhl)) || (rc = T.get(h, &ncols, np)) ||
        (rc = T.get(h, &nrows, np)) || (rc = T.get(h, &stripc, nvl)) || (rc = T.get(h, &stripn, nvl)) ||
        (rc = T.get(h, &stripmy, nvl)) || (rc = T.get(h, &psz, nvl + 1)) || (rc = T.get(h, &stripo, nvl + 1)) ||
        (rc = T.get(h, &hsz, nvl + 1)) || (rc = T.get(h, &hasho, nvl + 1)))
        return rc;
    HIPCHK(h, hipMemsetAsync(stripc, 0, 4 * nvl, h->stream));
    HIPCHK(h, hipMemsetAsync(stripn, 0, 4 * nvl, h->stream));
    if (np) {
        hipLaunchKernelGGL(k_rbounds, dim3(blocks_for(np)), dim3(kRB), 0, h->stream, (const RPoly *)d_rp, np, d_flag,
                           colx, rowy, ncols, nrows);
        HIPCHK(h, hipGetLastError());
        hipLaunchKernelGGL(k_rstrip<0>, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           P<float>(h->x), P<float>(h->y), (const float *)d_absx, d_flag, (const int *)ncols,
                           (const int *)nrows, (const double *)colx, stripc, (const uint64_t *)nullptr,
                           (const uint64_t *)nullptr, (float2 *)nullptr, (unsigned long long *)nullptr,
                           (uint32_t *)nullptr, stripn, stripmy);
        HIPCHK(h, hipGetLastError());
    }
    // pool sizes: every candidate plus the closing point; hashes in HBM for
    // strips beyond the LDS table
    {
        std::vector<uint32_t> sc(nvl);
        HIPCHK(h, hipMemcpyAsync(sc.data(), stripc, 4 * nvl, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        std::vector<uint64_t> po(nvl + 1), ho(nvl + 1);
        uint64_t tp = 0, th = 0;
        for (uint64_t s = 0; s < nvl; s++) {
            po[s] = tp;
            ho[s] = th;
            tp += sc[s] ? sc[s] + 1 : 0;
            if (sc[s] > kLdsCands) {
                uint64_t sl = 1;
                while (sl < 2ull * sc[s]) sl <<= 1;
                th += sl;
            }
        }
        po[nvl] = tp;
        ho[nvl] = th;
        HIPCHK(h, hipMemcpyAsync(stripo, po.data(), 8 * (nvl + 1), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(hasho, ho.data(), 8 * (nvl + 1), hipMemcpyHostToDevice, h->stream));
        float2 *pool;
        unsigned long long *hkeys;
        uint32_t *hidx;
        if ((rc = T.get(h, &pool, tp)) || (rc = T.get(h, &hkeys, th)) || (rc = T.get(h, &hidx, th))) return rc;
        if (np) {
            hipLaunchKernelGGL(k_rstrip<1>, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                               P<float>(h->x), P<float>(h->y), (const float *)d_absx, d_flag, (const int *)ncols,
                               (const int *)nrows, (const double *)colx, stripc, (const uint64_t *)stripo,
                               (const uint64_t *)hasho, pool, hkeys, hidx, stripn, stripmy);
            HIPCHK(h, hipGetLastError());
        }
        if (npix) {
            hipLaunchKernelGGL(k_rcells, dim3(grid_for(npix)), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                               (const uint32_t *)d_pixpoly, npix, (const uint32_t *)d_flag, (const int *)ncols,
                               (const int *)nrows, (const double *)rowy, (const uint64_t *)stripo,
                               (const uint32_t *)stripn, (const float *)stripmy, (const float2 *)pool,
                               P<uint8_t>(h->status), P<double>(h->areas));
            HIPCHK(h, hipGetLastError());
        }
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    (void)psz;
    (void)hsz;

    // ---- final arrays ----
    uint64_t *nseq, *nvnf, *seqf, *vnf;
    if ((rc = T.get(h, &nseq, np + 1)) || (rc = T.get(h, &nvnf, np + 1)) || (rc = T.get(h, &seqf, np + 1)) ||
        (rc = T.get(h, &vnf, np + 1)))
        return rc;
    hipLaunchKernelGGL(k_rcounts, dim3(blocks_for(np + 1)), dim3(kRB), 0, h->stream, (const RPoly *)d_rp, np, d_flag,
                       (const uint32_t *)seqo, (const uint32_t *)vlo, nseq, nvnf);
    HIPCHK(h, hipGetLastError());
    if ((rc = scan_excl(h, (const uint64_t *)nseq, seqf, np + 1)) ||
        (rc = scan_excl(h, (const uint64_t *)nvnf, vnf, np + 1)))
        return rc;
    uint64_t nseqs = 0, nvnodes = 0;
    if ((rc = read1(h, seqf + np, &nseqs)) || (rc = read1(h, vnf + np, &nvnodes))) return rc;
    if ((rc = ensure(h, h->pixoff, 4 * npo)) || (rc = ensure(h, h->seqs, 8 * nseqs)) ||
        (rc = ensure(h, h->voff, 4 * nvl)) || (rc = ensure(h, h->vnodes, 8 * nvnodes)))
        return rc;
    if (np) {
        hipLaunchKernelGGL(k_rpack, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                           (const uint32_t *)d_flag, (const uint64_t *)seqf, (const uint64_t *)vnf,
                           (const uint32_t *)seqo, (const uint2 *)seqtmp, (const uint32_t *)vlc,
                           (const uint32_t *)vlo, (const double *)vnv, P<farm_raster_hdr>(h->hdr),
                           P<uint8_t>(h->status), P<double>(h->areas), P<uint32_t>(h->pixoff), P<uint2>(h->seqs),
                           P<uint32_t>(h->voff), P<double>(h->vnodes));
        HIPCHK(h, hipGetLastError());
    }
    // ---- hierarchy ----
    h->has_layers = false;
    if (hier) {
        if ((rc = upload(h, h->layers, layers.data(), sizeof(farm_layer) * layers.size()))) return rc;
        if ((rc = ensure(h, h->lstatus, nls))) return rc;
        uint32_t *d_lpoly;
        unsigned long long *d_bmax;
        if ((rc = T.get(h, &d_lpoly, lpoly.size())) || (rc = T.get(h, &d_bmax, 1))) return rc;
        HIPCHK(h, hipMemcpyAsync(d_lpoly, lpoly.data(), 4 * lpoly.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemsetAsync(d_bmax, 0, 8, h->stream));
        if (nls) {
            hipLaunchKernelGGL(k_rmerge, dim3(grid_for(nls)), dim3(kRB), 0, h->stream,
                               (const farm_raster_hdr *)P<farm_raster_hdr>(h->hdr),
                               (const farm_layer *)P<farm_layer>(h->layers), (const uint32_t *)d_lpoly,
                               (uint64_t)layers.size(), nls, (const uint8_t *)P<uint8_t>(h->status),
                               (const double *)P<double>(h->areas), P<uint8_t>(h->lstatus));
            HIPCHK(h, hipGetLastError());
        }
        if (np) {
            hipLaunchKernelGGL(k_rborder_max, dim3((unsigned)np), dim3(kRB), 0, h->stream, (const RPoly *)d_rp,
                               (const uint8_t *)P<uint8_t>(h->status), d_bmax);
            HIPCHK(h, hipGetLastError());
        }
        unsigned long long bmax = 0;
        if ((rc = read1(h, (const unsigned long long *)d_bmax, &bmax))) return rc;
        h->has_layers = true;
        h->max_layer_pix = max_layer_pix;
        h->max_border_pix = bmax;
    }
    std::vector<uint32_t> flags(np);
    if (np) HIPCHK(h, hipMemcpyAsync(flags.data(), d_flag, 4 * np, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int64_t nrej = 0, first = -1;
    for (int64_t i = 0; i < np; i++)
        if (flags[i]) {
            if (flags[i] & kRejUnsupported)
                return fail(h, FARM_EINVAL, "polygon %lld: the raster loops overran the grid", (long long)i);
            if (first < 0) first = i;
            nrej++;
        }
    h->cc = FARM_CATEGORY_COUNT;
    h->areas_ready = true;
    h->npolys = np;
    h->nstatus = npix;
    h->nverts = nvt;
    h->nseqs = nseqs;
    h->npixoff = npo;
    h->nvoff = nvl;
    h->nvnodes = nvnodes;
    h->nlayers = hier ? layers.size() : 0;
    h->nlstatus = hier ? nls : 0;
    if ((rc = finish_set(h))) return rc;
    HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev[1]));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    if (info) {
        memset(info, 0, sizeof *info);
        info->build_ms = ms;
        info->nrejected = nrej;
        info->first_rejected = first;
        info->nverts = nvt;
        info->nstatus = npix;
        info->npixoff = npo;
        info->nseqs = nseqs;
        info->nvoff = nvl;
        info->nvnodes = nvnodes;
        info->nlayers = h->nlayers;
        info->nlstatus = h->nlstatus;
    }
    return FARM_OK;
}

}  // namespace
}  // namespace farm

extern "C" int farm_rasterize(farm_handle h, const farm_ringset *rs, int vpr, int nlow, farm_raster_info *info) {
    if (!h || !rs || rs->npolys < 0 || vpr <= 0 || nlow < 0 || (rs->npolys && (!rs->vert_off || !rs->x || !rs->y)))
        return fail(h, FARM_EINVAL, "farm_rasterize: bad arguments");
    const int64_t np = rs->npolys;
    const uint64_t vbase = np ? rs->vert_off[0] : 0;
    for (int64_t i = 0; i < np; i++) {
        const uint64_t a = rs->vert_off[i], b = rs->vert_off[i + 1];
        if (b < a || b - a < 2 || b - a >= (1ull << 31))
            return fail(h, FARM_EINVAL, "polygon %lld: ring of %lld vertices", (long long)i, (long long)(b - a));
    }
    const uint64_t nvt = np ? rs->vert_off[np] - vbase : 0;
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (DevBuf *b : {&h->x, &h->y})
        if ((rc = ensure(h, *b, 4 * (nvt + kVertexPad)))) return rc;
    if ((rc = upload(h, h->x, rs->x + vbase, 4 * nvt))) return rc;
    if ((rc = upload(h, h->y, rs->y + vbase, 4 * nvt))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double up_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<uint64_t> voff(np + 1);
    for (int64_t i = 0; i <= np; i++) voff[i] = np ? rs->vert_off[i] - vbase : 0;
    rc = rasterize_uploaded(h, np, voff, nvt, vpr, nlow, info);
    if (!rc && info) info->upload_ms = up_ms;
    return rc;
}

// The host set's rings go to the device through two pinned staging buffers:
// host threads write chunk k + 1 as closed SoA rings while chunk k's DMA
// runs, then the set is rasterized there.
extern "C" int farm_hostset_rasterize_device(farm_handle h, const farm_hostset *s, int vpr, int nlow, int nthreads,
                                             farm_raster_info *info) {
    if (!h || !s || vpr <= 0 || nlow < 0) return fail(h, FARM_EINVAL, "farm_hostset_rasterize_device: bad arguments");
    const int64_t np = farm_hostset_size(s);
    std::vector<uint64_t> voff(np + 1, 0);
    std::vector<const float *> rings(np);
    for (int64_t i = 0; i < np; i++) {
        int64_t nv = 0;
        if (farm_hostset_ring(s, i, &rings[i], &nv) || nv + 1 >= (1ll << 31))
            return fail(h, FARM_EINVAL, "polygon %lld: bad ring", (long long)i);
        voff[i + 1] = voff[i] + (uint64_t)nv + 1;
    }
    const uint64_t nvt = voff[np];
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (DevBuf *b : {&h->x, &h->y})
        if ((rc = ensure(h, *b, 4 * (nvt + kVertexPad)))) return rc;
    const uint64_t kStage = h->stage_verts;  // vertices per staging buffer (x then y)
    for (int k = 0; k < 2; k++) {
        if (!h->stage[k]) HIPCHK(h, hipHostMalloc(&h->stage[k], 8 * kStage, hipHostMallocDefault));
        if (!h->stage_ev[k]) HIPCHK(h, hipEventCreateWithFlags(&h->stage_ev[k], hipEventDisableTiming));
    }
    const int nth = std::max(1, nthreads);
    int64_t i0 = 0;
    int k = 0;
    bool pending[2] = {false, false};
    while (i0 < np) {
        // the next chunk: whole rings up to kStage vertices (a longer ring alone)
        int64_t i1 = i0 + 1;
        while (i1 < np && voff[i1 + 1] - voff[i0] <= kStage) i1++;
        const uint64_t v0 = voff[i0], nvc = voff[i1] - v0;
        float *sx = (float *)h->stage[k], *sy = sx + kStage;
        DevBuf big;  // a ring longer than a staging buffer goes through a pinned buffer of its own
        if (nvc > kStage) {
            HIPCHK(h, hipHostMalloc(&big.p, 8 * nvc, hipHostMallocDefault));
            sx = (float *)big.p;
            sy = sx + nvc;
        } else if (pending[k]) {
            HIPCHK(h, hipEventSynchronize(h->stage_ev[k]));  // buffer k's last DMA is done
        }
        auto work = [&](int t) {
            for (int64_t i = i0 + t; i < i1; i += nth) {
                const uint64_t o = voff[i] - v0, n = voff[i + 1] - voff[i] - 1;
                const float *r = rings[i];
                for (uint64_t v = 0; v < n; v++) {
                    sx[o + v] = r[2 * v];
                    sy[o + v] = r[2 * v + 1];
                }
                sx[o + n] = r[0];  // the decoder's repeated first vertex
                sy[o + n] = r[1];
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nth && t < i1 - i0; t++) th.emplace_back(work, t);
        work(0);
        for (auto &q : th) q.join();
        HIPCHK(h, hipMemcpyAsync(P<float>(h->x) + v0, sx, 4 * nvc, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(P<float>(h->y) + v0, sy, 4 * nvc, hipMemcpyHostToDevice, h->stream));
        if (big.p) {
            HIPCHK(h, hipStreamSynchronize(h->stream));
            HIPCHK(h, hipHostFree(big.p));
        } else {
            HIPCHK(h, hipEventRecord(h->stage_ev[k], h->stream));
            pending[k] = true;
            k ^= 1;
        }
        i0 = i1;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double up_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    rc = rasterize_uploaded(h, np, voff, nvt, vpr, nlow, info);
    if (!rc && info) info->upload_ms = up_ms;
    return rc;
}

extern "C" int farm_set_staging(farm_handle h, uint64_t vertices) {
    if (!h || vertices < 1024 || vertices > (1ull << 28)) return fail(h, FARM_EINVAL, "staging of %llu vertices",
                                                                      (unsigned long long)vertices);
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 2; k++)
        if (h->stage[k]) {
            HIPCHK(h, hipHostFree(h->stage[k]));
            h->stage[k] = nullptr;
        }
    h->stage_verts = vertices;
    return FARM_OK;
}

extern "C" int farm_set_categories(farm_handle h, int category_bits) {
    if (!h || category_bits < 2 || category_bits > 8) return fail(h, FARM_EINVAL, "-w %d: bits in 2..8", category_bits);
    if (h->npolys <= 0 || !h->areas_ready) return fail(h, FARM_ESTATE, "no device set with per-pixel areas");
    HIPCHK(h, hipSetDevice(h->device));
    const int cc = 1 << category_bits;
    if (cc != FARM_CATEGORY_COUNT) {
        int rc;
        if ((rc = ensure(h, h->fill, h->nstatus))) return rc;
        hipLaunchKernelGGL(k_rfill, dim3((unsigned)h->npolys), dim3(kRB), 0, h->stream,
                           (const farm_raster_hdr *)P<farm_raster_hdr>(h->hdr), (const uint8_t *)P<uint8_t>(h->status),
                           (const double *)P<double>(h->areas), cc, P<uint8_t>(h->fill));
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    h->cc = cc;
    h->set_gen++;
    return FARM_OK;
}

extern "C" int farm_download_polyset(farm_handle h, farm_polyset_buf *o) {
    if (!h || !o) return fail(h, FARM_EINVAL, "farm_download_polyset: bad arguments");
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygon set on the device");
    HIPCHK(h, hipSetDevice(h->device));
    auto get = [&](void *dst, DevBuf &b, size_t bytes) -> int {
        if (dst && bytes) HIPCHK(h, hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToHost, h->stream));
        return FARM_OK;
    };
    int rc;
    if ((rc = get(o->hdr, h->hdr, sizeof(farm_raster_hdr) * h->npolys)) || (rc = get(o->status, h->status, h->nstatus)) ||
        (rc = get(o->pixoff, h->pixoff, 4 * h->npixoff)) || (rc = get(o->seqs, h->seqs, 8 * h->nseqs)) ||
        (rc = get(o->voff, h->voff, 4 * h->nvoff)) || (rc = get(o->vnodes, h->vnodes, 8 * h->nvnodes)))
        return rc;
    if (h->has_layers &&
        ((rc = get(o->layers, h->layers, sizeof(farm_layer) * h->nlayers)) ||
         (rc = get(o->lstatus, h->lstatus, h->nlstatus))))
        return rc;
    if (h->areas_ready && (rc = get(o->areas, h->areas, 8 * h->nstatus))) return rc;
    if (h->cc != FARM_CATEGORY_COUNT && (rc = get(o->fill, h->fill, h->nstatus))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}
