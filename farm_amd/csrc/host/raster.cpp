// Host builder of the IDEAL raster model for one polygon.
//
// Semantics follow the reference CPU raster builder exactly (same pixel grid,
// same BORDER marking walk, same edge sequences, same scanline parity, same
// clipped-area fill classes):
//   grid sizing          MyRaster::init_raster(int)      MyRaster.cpp:8-33
//   edge walk            Ideal::evaluate_edges           Ideal.cpp:252-538
//   edge sequences       Ideal::process_crosses          Ideal.cpp:90-183
//   grid-line nodes      Ideal::process_intersection     Ideal.cpp:185-238
//   CSR fill-in          Ideal::process_pixels_null      Ideal.cpp:35-45
//   IN/OUT               Ideal::scanline_reandering      Ideal.cpp:540-573
//   fill classes         Ideal::calculate_fullness       Ideal.cpp:820-918
//                        (intersectionDualX/Y Ideal.cpp:575-818,
//                         computePolygonArea geometry_computation.h:275-297,
//                         encodePixelArea Ideal.cpp:64-88)
// The data structures are this project's own: events are bucketed per pixel
// with a stable counting sort instead of per-pixel std::map vectors, the
// clipped strips are built once per column from the edges that reach it, and
// the tolerant point dedup runs through a hash set whenever tolerant and
// exact float equality provably coincide.
#include "raster.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace farm {

namespace {

constexpr int kCC = 16;  // fill categories (MyRaster.h:33)
constexpr uint8_t kBorderRaw = 1;
constexpr uint8_t kInRaw = 2;

struct P2 {
    float x, y;
};

inline int dbl2int(double v) {  // util.h:98-106
    int i = (int)v;
    if (std::abs(1.0 * (i + 1) - v) < 0.00000001) i++;
    return i;
}

// Point::operator== / != (Point.h:32-33): float difference within 1e-9
inline bool same_pt(const P2 &a, const P2 &b) {
    float dx = a.x - b.x, dy = a.y - b.y;
    return std::fabs(dx) < 1e-9 && std::fabs(dy) < 1e-9;
}

struct Event {
    int pix;
    int edge;
    int leave;  // 1 LEAVE, 0 ENTER
};
struct Node {
    int line;
    double v;
};

}  // namespace

struct RasterScratch {
    std::vector<Event> ev, ev_sorted;
    std::vector<int> cnt;
    std::vector<Node> hn, vn;
    std::vector<P2> strip, cell;
    std::vector<uint64_t> hash;
    std::vector<int> strip_start;
    std::vector<P2> strips;
    std::vector<int> strip_col;
    std::vector<int> bcnt, bidx, brng;  // edges bucketed by column / row
};

namespace {

// Edges k = 0..n-2 of the polyline V, bucketed by the cells of one axis they
// can touch (cell c spans [lo + c*step, lo + (c+1)*step]), in polyline order.
// An edge spanning [mn, mx] on the axis adds a point to cell c only if a cell
// bound, rounded to float, lies in [mn, mx] or an end point lies in the cell:
// the lower bound of cell c + 1 counts for cell c, and a float rounding moves
// a bound by at most |x| * 2^-24.  The margin m covers that rounding, the
// accumulated cell bounds (one double rounding per cell) and the rounding of
// the bucket arithmetic, so the cells [floor((mn - m - lo) / step),
// floor((mx + m - lo) / step)] hold every cell the clip tests can reach.
void bucket_edges(const float *V, int n, int axis, double lo, double step, int ncell, std::vector<int> &cnt,
                  std::vector<int> &idx, std::vector<int> &rng) {
    cnt.assign((size_t)ncell + 1, 0);
    rng.resize(2 * (size_t)std::max(n - 1, 0));
    const double inv = 1.0 / step;
    double amax = std::max(std::fabs(lo), std::fabs(lo + (ncell + 1) * step));
    for (int k = 0; k < n; k++) amax = std::max(amax, (double)std::fabs(V[2 * k + axis]));
    const double m = amax * (0x1p-22 + (ncell + 8) * 0x1p-50) + step * 0x1p-20;
    float a = V[axis];
    for (int k = 0; k + 1 < n; k++) {
        const float b = V[2 * (k + 1) + axis];
        const double mn = std::min(a, b), mx = std::max(a, b);
        int c0 = (int)std::floor((mn - m - lo) * inv);
        int c1 = (int)std::floor((mx + m - lo) * inv);
        c0 = std::max(c0, 0);
        c1 = std::min(c1, ncell - 1);
        rng[2 * k] = c0;
        rng[2 * k + 1] = c1;
        for (int c = c0; c <= c1; c++) cnt[c + 1]++;
        a = b;
    }
    for (int c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    idx.resize(cnt[ncell]);
    std::vector<int> &pos = rng;  // reuse: cursors live past the ranges
    const size_t base = rng.size();
    pos.resize(base + ncell);
    for (int c = 0; c < ncell; c++) pos[base + c] = cnt[c];
    for (int k = 0; k + 1 < n; k++)
        for (int c = rng[2 * k]; c <= rng[2 * k + 1]; c++) idx[pos[base + c]++] = k;
}

}  // namespace

RasterScratch *raster_scratch_new() { return new RasterScratch(); }
void raster_scratch_free(RasterScratch *s) { delete s; }

RasterGrid size_grid(const double mbb[4], int num_pixels) {
    RasterGrid g;
    g.low[0] = mbb[0];
    g.low[1] = mbb[1];
    g.high[0] = mbb[2];
    g.high[1] = mbb[3];
    const double w = mbb[2] - mbb[0], h = mbb[3] - mbb[1];
    const double aspect = std::abs(h / w);
    g.dimx = (int)std::pow(num_pixels / aspect, 0.5);
    g.dimy = (int)(g.dimx * aspect);
    if (g.dimx == 0) g.dimx = 1;
    if (g.dimy == 0) g.dimy = 1;
    g.step[0] = w / g.dimx;
    g.step[1] = h / g.dimy;
    if (g.step[0] < 0.00001) {
        g.step[0] = 0.00001;
        g.dimx = (int)(w / g.step[0] + 1);
    }
    if (g.step[1] < 0.00001) {
        g.step[1] = 0.00001;
        g.dimy = (int)(h / g.step[1] + 1);
    }
    return g;
}

namespace {

// Point dedup of TempPolygon::addPoint (MyPolygon.h:49-57).  When every
// coordinate the strip can hold has |v| >= 2^-6, two different floats on one
// axis differ by at least 2^-29 > 1e-9 (same sign: one ulp at 2^-6; opposite
// signs: 2^-5), so tolerant equality is exact equality and a hash set answers
// it.
struct Dedup {
    bool exact;
    std::vector<uint64_t> *tab;
    uint64_t mask = 0;
    size_t n = 0;
    void reset(size_t cap) {
        if (!exact) return;
        size_t sz = 64;
        while (sz < cap * 2) sz <<= 1;
        tab->assign(sz, ~0ull);
        mask = sz - 1;
        n = 0;
    }
    static uint64_t key(const P2 &p) {
        uint32_t a, b;
        std::memcpy(&a, &p.x, 4);
        std::memcpy(&b, &p.y, 4);
        return ((uint64_t)a << 32) | b;
    }
    // returns true if p was not present (and records it)
    bool insert(std::vector<P2> &v, const P2 &p) {
        if (!exact) {
            for (const P2 &q : v)
                if (same_pt(q, p)) return false;
            v.push_back(p);
            return true;
        }
        uint64_t k = key(p);
        uint64_t hsh = k * 0x9E3779B97F4A7C15ull;
        size_t i = (hsh >> 20) & mask;
        while ((*tab)[i] != ~0ull) {
            if ((*tab)[i] == k) return false;
            i = (i + 1) & mask;
        }
        if ((n + 1) * 2 > tab->size()) {  // grow and rehash
            std::vector<uint64_t> old;
            old.swap(*tab);
            tab->assign(old.size() * 2, ~0ull);
            mask = tab->size() - 1;
            for (uint64_t o : old) {
                if (o == ~0ull) continue;
                size_t j = ((o * 0x9E3779B97F4A7C15ull) >> 20) & mask;
                while ((*tab)[j] != ~0ull) j = (j + 1) & mask;
                (*tab)[j] = o;
            }
            i = (hsh >> 20) & mask;
            while ((*tab)[i] != ~0ull) i = (i + 1) & mask;
        }
        (*tab)[i] = k;
        n++;
        v.push_back(p);
        return true;
    }
};

inline bool within(double lo, double hi, double v) { return !(v < lo || v > hi); }

}  // namespace

// the raster on a given grid (out.g)
static void build_raster_impl(const float *ring, int nv, RasterOut &out, RasterScratch *S) {
    out.ok = true;
    const RasterGrid &g = out.g;
    const int dimx = g.dimx, dimy = g.dimy;
    const int npix = dimx * dimy;
    const double x0 = g.low[0], y0 = g.low[1];
    const double sx = g.step[0], sy = g.step[1];
    const P2 *p = reinterpret_cast<const P2 *>(ring);

    out.status.assign(npix, 0);
    out.areas.assign(npix, 0.0);
    out.pixoff.assign((size_t)npix + 1, 0);
    out.voff.assign((size_t)dimx + 2, 0);
    out.hoff.assign((size_t)dimy + 2, 0);
    out.seqs.clear();
    out.vnodes.clear();
    out.hnodes.clear();

    auto &ev = S->ev;
    auto &hn = S->hn;
    auto &vn = S->vn;
    ev.clear();
    hn.clear();
    vn.clear();
    uint8_t *st = out.status.data();

    auto pid = [&](int x, int y) -> int {
        if (x < 0 || x >= dimx || y < 0 || y >= dimy) {
            out.ok = false;
            x = std::min(std::max(x, 0), dimx - 1);
            y = std::min(std::max(y, 0), dimy - 1);
        }
        return y * dimx + x;
    };
    // one pixel-to-pixel step of the walk: LEAVE the old pixel, ENTER the new
    auto step_to = [&](int ox, int oy, int nx, int ny, int e) {
        int a = pid(ox, oy), b = pid(nx, ny);
        st[a] = kBorderRaw;
        ev.push_back({a, e, 1});
        ev.push_back({b, e, 0});
        st[b] = kBorderRaw;
    };

    // ---- edge walk (BORDER pixels, crossings, grid-line nodes) ----
    for (int e = 0; e < nv - 1 && out.ok; e++) {
        const double ax = p[e].x, ay = p[e].y, bx = p[e + 1].x, by = p[e + 1].y;
        int cx0 = dbl2int((ax - x0) / sx), cx1 = dbl2int((bx - x0) / sx);
        int cy0 = dbl2int((ay - y0) / sy), cy1 = dbl2int((by - y0) / sy);
        if (cx0 == dimx) cx0--;
        if (cx1 == dimx) cx1--;
        if (cy0 == dimy) cy0--;
        if (cy1 == dimy) cy1--;
        if (cx0 >= dimx || cx1 >= dimx || cy0 >= dimy || cy1 >= dimy) {
            out.ok = false;
            break;
        }
        st[pid(cx0, cy0)] = kBorderRaw;
        st[pid(cx1, cy1)] = kBorderRaw;
        if (cx0 == cx1 && cy0 == cy1) continue;

        if (ay == by) {  // horizontal edge
            const int d = cx0 < cx1 ? 1 : -1;
            for (int x = cx0; x != cx1; x += d) {
                vn.push_back({d > 0 ? x + 1 : x, ay});
                step_to(x, cy0, x + d, cy0, e);
            }
            continue;
        }
        if (ax == bx) {  // vertical edge
            const int d = cy0 < cy1 ? 1 : -1;
            for (int y = cy0; y != cy1; y += d) {
                hn.push_back({d > 0 ? y + 1 : y, ax});
                step_to(cx0, y, cx0, y + d, e);
            }
            continue;
        }
        // general edge: follow the line y = k*x + c through the grid
        const double k = (ay - by) / (ax - bx);
        const double c = (ax * by - bx * ay) / (ax - bx);
        const int dx = cx0 < cx1 ? 1 : -1, dy = cy0 < cy1 ? 1 : -1;
        const int ylo = std::min(cy0, cy1), yhi = std::max(cy0, cy1);
        const int xlo = std::min(cx0, cx1), xhi = std::max(cx0, cx1);
        int x = cx0, y = cy0;
        int64_t guard = 4LL * ((int64_t)dimx + dimy) + 16;
        while ((x != cx1 || y != cy1) && out.ok) {
            if (guard-- <= 0) {
                out.ok = false;
                break;
            }
            bool moved = false;
            if (x != cx1) {
                const double xv = dx > 0 ? ((double)x + 1) * sx + x0 : (double)x * sx + x0;
                const double yv = xv * k + c;
                const double q = (yv - y0) / sy;
                if (std::abs(q - std::round(q)) < 1e-9) {
                    // the line passes a grid corner: step diagonally; only the
                    // rightward walks record a vertical node (Ideal.cpp:397-423).
                    // An edge that stays in one pixel row matches none of the
                    // reference's four diagonal cases and its walk never ends.
                    if (cy0 == cy1) {
                        out.ok = false;
                        break;
                    }
                    if (dx > 0) vn.push_back({dy > 0 ? x + 1 : x, yv});
                    step_to(x, y, x + dx, y + dy, e);
                    x += dx;
                    y += dy;
                    continue;
                }
                int ry = (int)((yv - y0) / sy);
                ry = std::min(std::max(ry, ylo), yhi);
                if (ry == y) {
                    moved = true;
                    vn.push_back({dx > 0 ? x + 1 : x, yv});
                    step_to(x, y, x + dx, y, e);
                    x += dx;
                }
            }
            if (y != cy1) {
                const double yv = dy > 0 ? (y + 1) * sy + y0 : y * sy + y0;
                const double xv = (yv - c) / k;
                int rx = (int)((xv - x0) / sx);
                rx = std::min(std::max(rx, xlo), xhi);
                if (rx == x) {
                    moved = true;
                    hn.push_back({dy > 0 ? y + 1 : y, xv});
                    step_to(x, y, x, y + dy, e);
                    y += dy;
                }
            }
            if (!moved) out.ok = false;
        }
    }
    if (!out.ok) return;

    // ---- edge sequences per pixel ----
    auto &cnt = S->cnt;
    if (ev.empty()) {
        // The whole ring lies in one pixel.  The reference records
        // (0, num_vertices) here, one edge more than the ring has, and its
        // refinement reads one vertex past the array; the real edges are kept.
        out.seqs.push_back(0);
        out.seqs.push_back((uint32_t)(nv - 1));
        out.pixoff[0] = 0;
    } else {
        cnt.assign((size_t)npix + 1, 0);
        for (const Event &v : ev) cnt[v.pix + 1]++;
        for (int i = 0; i < npix; i++) cnt[i + 1] += cnt[i];
        auto &sorted = S->ev_sorted;
        sorted.resize(ev.size());
        {
            std::vector<int> pos(cnt.begin(), cnt.end() - 1);
            for (const Event &v : ev) sorted[pos[v.pix]++] = v;
        }
        uint32_t nseq = 0;
        for (int pix = 0; pix < npix && out.ok; pix++) {
            int b = cnt[pix], m = cnt[pix + 1] - b;
            if (m == 0) continue;
            const Event *E = &sorted[b];
            out.pixoff[pix] = nseq;
            // an odd list is closed with the opposite type of its last entry
            auto type_at = [&](int i) { return i < m ? E[i].leave : !E[m - 1].leave; };
            auto edge_at = [&](int i) { return i < m ? E[i].edge : E[m - 1].edge; };
            const int mm = m + (m & 1);
            auto push = [&](int s, int t) {
                if (t - s + 1 <= 0) out.ok = false;
                out.seqs.push_back((uint32_t)s);
                out.seqs.push_back((uint32_t)(t - s + 1));
                nseq++;
            };
            int i = 0, last = mm - 1;
            if (type_at(0) == 1) {  // the ring starts inside this pixel
                if (type_at(last) != 0) out.ok = false;
                push(0, edge_at(0));
                push(edge_at(last), nv - 2);
                i = 1;
                last--;
            }
            for (; i <= last; i++) {
                if (type_at(i) != 0) out.ok = false;
                if (i == last || type_at(i + 1) == 0) {
                    push(edge_at(i), edge_at(i));
                } else {
                    push(edge_at(i), edge_at(i + 1));
                    i++;
                }
            }
        }
        if (nseq >= 65536) out.ok = false;  // Ideal.cpp:32
        // a BORDER pixel that got no crossing has no defined sequence range
        for (int pix = 0; pix < npix; pix++)
            if (st[pix] == kBorderRaw && cnt[pix + 1] == cnt[pix]) out.ok = false;
    }
    if (!out.ok) return;
    const uint32_t nseq_total = (uint32_t)(out.seqs.size() / 2);
    out.pixoff[npix] = nseq_total;
    for (int i = npix - 1; i >= 0; i--)
        if (st[i] != kBorderRaw) out.pixoff[i] = out.pixoff[i + 1];

    // ---- grid-line nodes: bucket per line, sort each line ----
    auto fill_lines = [&](std::vector<Node> &nodes, int nlines, std::vector<uint32_t> &off, std::vector<double> &vals,
                          bool vertical) {
        vals.resize(nodes.size());
        std::vector<int> c(nlines + 2, 0);
        for (const Node &n : nodes) c[n.line + 1]++;
        for (int i = 0; i <= nlines; i++) c[i + 1] += c[i];
        std::vector<int> pos(c.begin(), c.end() - 1);
        for (const Node &n : nodes) vals[pos[n.line]++] = n.v;
        // lines without nodes keep offset 0 (Grid_line.cpp:3-7)
        for (int l = 0; l <= nlines; l++) {
            if (c[l + 1] > c[l]) {
                off[l] = (uint32_t)c[l];
                std::sort(vals.begin() + c[l], vals.begin() + c[l + 1]);
            }
        }
        if (vertical) {
            off[nlines + 1] = (uint32_t)nodes.size();
            off[nlines] = (uint32_t)nodes.size();
        } else {
            off[nlines] = (uint32_t)nodes.size();
        }
    };
    fill_lines(hn, dimy, out.hoff, out.hnodes, false);
    fill_lines(vn, dimx, out.voff, out.vnodes, true);

    // ---- scanline parity: IN/OUT for rows >= 1 ----
    for (int y = 1; y < dimy; y++) {
        bool inside = false;
        uint32_t i = out.hoff[y], j = out.hoff[y + 1];
        for (int x = 0; x < dimx; x++) {
            uint8_t &s = st[y * dimx + x];
            if (s != kBorderRaw) {
                s = inside ? kInRaw : 0;
                continue;
            }
            int pass = 0;
            const double right = x0 + sx * (x + 1);
            while (i < j && out.hnodes[i] <= right) {
                pass++;
                i++;
            }
            if (pass & 1) inside = !inside;
        }
    }

    // ---- fill classes: clip to each column strip, then to each pixel ----
    const double pix_area = sx * sy;
    // every clip point lies in the grid box (one cell of slack past its high
    // side); when neither axis of it reaches into [-0.016, 0.016], distinct
    // floats are more than 1e-9 apart and the tolerant dedup is an exact one
    auto off_zero = [](double lo, double hi) { return lo > 0.0160 || hi < -0.0160; };
    const bool exact_dedup = off_zero(x0, x0 + (dimx + 1) * sx) && off_zero(y0, y0 + (dimy + 1) * sy);
    Dedup dd{exact_dedup, &S->hash};
    auto &strips = S->strips;
    auto &sstart = S->strip_start;
    auto &scol = S->strip_col;
    strips.clear();
    sstart.clear();
    scol.clear();

    {
        double Xi = x0, Xi1 = Xi + sx;
        const double kx = Xi + dimx * sx;
        int col = 0;
        const int ncol = dimx + 2;
        bucket_edges(ring, nv, 0, x0, sx, ncol, S->bcnt, S->bidx, S->brng);
        while (Xi1 < kx + 1e-9) {
            const size_t base = strips.size();
            S->strip.clear();
            dd.reset(64);
            const float fXi = (float)Xi, fXi1 = (float)Xi1;
            // only the edges that can reach this column, in ring order
            const bool bucketed = col < ncol;
            const int kb = bucketed ? S->bcnt[col] : 0, ke = bucketed ? S->bcnt[col + 1] : nv - 1;
            for (int kk = kb; kk < ke; kk++) {
                const int e = bucketed ? S->bidx[kk] : kk;
                const P2 A = p[e], B = p[e + 1];
                double yi, yi1;
                if (A.x == B.x) {
                    yi = std::min(A.y, B.y);
                    yi1 = yi;
                } else {
                    const float den = B.x - A.x;
                    const double slope = ((double)B.y - A.y) / den;
                    yi = slope * (Xi - B.x) + B.y;
                    yi1 = slope * (Xi1 - B.x) + B.y;
                }
                const P2 q0 = {fXi, (float)yi}, q1 = {fXi1, (float)yi1};
                const float mny = std::min(A.y, B.y), mxy = std::max(A.y, B.y);
                const float mnx = std::min(A.x, B.x), mxx = std::max(A.x, B.x);
                const bool in0 = within(mny, mxy, q0.y) && within(mnx, mxx, q0.x);
                const bool in1 = within(mny, mxy, q1.y) && within(mnx, mxx, q1.x);
                if (within(Xi, Xi1, A.x)) dd.insert(S->strip, A);
                if (A.x < B.x) {
                    if (in0) dd.insert(S->strip, q0);
                    if (in1) dd.insert(S->strip, q1);
                } else {
                    if (in1) dd.insert(S->strip, q1);
                    if (in0) dd.insert(S->strip, q0);
                }
                if (within(Xi, Xi1, B.x)) dd.insert(S->strip, B);
            }
            if (!S->strip.empty()) {
                S->strip.push_back(S->strip.front());
                sstart.push_back((int)base);
                scol.push_back(col);
                strips.insert(strips.end(), S->strip.begin(), S->strip.end());
            }
            Xi += sx;
            Xi1 = Xi + sx;
            col++;
        }
        sstart.push_back((int)strips.size());
    }

    const double ky = y0 + dimy * sy;
    auto &cell = S->cell;
    const int nrow = dimy + 2;
    for (size_t si = 0; si + 1 < sstart.size(); si++) {
        const P2 *V = strips.data() + sstart[si];
        const int nvs = sstart[si + 1] - sstart[si];
        const int col = scol[si];
        bucket_edges(reinterpret_cast<const float *>(V), nvs, 1, y0, sy, nrow, S->bcnt, S->bidx, S->brng);
        double Yi = y0, Yi1 = Yi + sy;
        int row = 0;
        while (Yi1 < ky + 1e-9) {
            cell.clear();
            const bool bucketed = row < nrow;
            const int kb = bucketed ? S->bcnt[row] : 0, ke = bucketed ? S->bcnt[row + 1] : nvs - 1;
            for (int kk = kb; kk < ke; kk++) {
                const int k = bucketed ? S->bidx[kk] : kk;
                const P2 A = V[k], B = V[k + 1];
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
                if (within(Yi, Yi1, A.y)) cell.push_back(A);
                if (skip) continue;
                const P2 q0 = {(float)xi, (float)Yi}, q1 = {(float)xi1, (float)Yi1};
                const float mny = std::min(A.y, B.y), mxy = std::max(A.y, B.y);
                const float mnx = std::min(A.x, B.x), mxx = std::max(A.x, B.x);
                const bool in0 = ok0 && within(mnx, mxx, q0.x) && within(mny, mxy, q0.y);
                const bool in1 = ok1 && within(mnx, mxx, q1.x) && within(mny, mxy, q1.y);
                if (A.y < B.y) {
                    if (in0) cell.push_back(q0);
                    if (in1) cell.push_back(q1);
                } else {
                    if (in1) cell.push_back(q1);
                    if (in0) cell.push_back(q0);
                }
            }
            if (!cell.empty() && !same_pt(cell.front(), cell.back())) cell.push_back(cell.front());
            double area = 0.0;
            if (cell.size() >= 3) {
                for (size_t k = 0; k + 1 < cell.size(); k++)
                    area += ((double)cell[k].x * cell[k + 1].y - (double)cell[k + 1].x * cell[k].y);
                area = std::abs(area) / 2.0;
            }
            // encodePixelArea
            int cls;
            const double ratio = area / pix_area;
            if (std::fabs(ratio - 1.0) < 1e-9) cls = kCC - 1;
            else if (std::fabs(ratio) < 1e-9) cls = 0;
            else {
                cls = (int)std::ceil(ratio * (kCC - 2));
                if (cls >= kCC) cls = kCC - 1;
                cls = (uint8_t)cls;
            }
            if (col >= dimx || row >= dimy) {
                out.ok = false;
                return;
            }
            const int id = row * dimx + col;
            if (st[id] == kBorderRaw) {
                st[id] = (uint8_t)std::max(1, std::min(cls, kCC - 2));
                out.areas[id] = area;
            } else if (st[id] == kInRaw) {
                st[id] = kCC - 1;
                out.areas[id] = pix_area;
            } else {
                st[id] = 0;
                out.areas[id] = 0.0;
            }
            Yi += sy;
            Yi1 = Yi + sy;
            row++;
        }
    }
}

static void make_empty(RasterOut &out);

void build_raster(const float *ring, int nv, const double mbb[4], int vpr, RasterOut &out, RasterScratch *S) {
    out.g = size_grid(mbb, nv / vpr);
    build_raster_impl(ring, nv, out, S);
    if (!out.ok) make_empty(out);
}

namespace {

// UniversalGrid::configure(25) (UniversalGrid.h:29-35): 360 / 2^25 rounded
// to two significant digits (roundToSignificantDigits, util.h:51-63)
double universal_step() {
    const double s = 360.0 / std::pow(2, 25);
    const double d = std::ceil(std::log10(s));
    const int power = 2 - static_cast<int>(d);
    const double mag = std::pow(10.0, power);
    return std::round(s * mag) / mag;
}

// MyRaster::grid_align (MyRaster.cpp:444-491): snap both steps to the
// nearer of the two bracketing universal steps 1.1e-5 * 2^k, take the
// smaller as the square pixel size, floor/ceil the box to it
bool align_grid(RasterGrid &g) {
    const double us = universal_step();
    double x = us, y = us;
    bool fx = false, fy = false;
    if (g.step[0] <= x) g.step[0] = x, fx = true;
    if (g.step[1] <= y) g.step[1] = y, fy = true;
    for (int i = 25; i > 0; i--) {
        if (!fx && g.step[0] > x && g.step[0] < x * 2) {
            fx = true;
            g.step[0] = std::abs(g.step[0] - x) < std::abs(g.step[0] - x * 2) ? x : x * 2;
        }
        if (!fy && g.step[1] > y && g.step[1] < y * 2) {
            fy = true;
            g.step[1] = std::abs(g.step[1] - y) < std::abs(g.step[1] - y * 2) ? y : y * 2;
        }
        if (fx && fy) break;
        x *= 2;
        y *= 2;
    }
    if (!fx || !fy) return false;
    const double s = std::min(g.step[0], g.step[1]);
    g.step[0] = g.step[1] = s;
    g.low[0] = std::floor(g.low[0] / s) * s;
    g.low[1] = std::floor(g.low[1] / s) * s;
    g.high[0] = std::ceil(g.high[0] / s) * s;
    g.high[1] = std::ceil(g.high[1] / s) * s;
    g.dimx = static_cast<int>(std::round((g.high[0] - g.low[0]) / s));
    g.dimy = static_cast<int>(std::round((g.high[1] - g.low[1]) / s));
    return g.dimx > 0 && g.dimy > 0;
}

inline int grid_off(double v, double lo, double step, int dim) {
    return std::min(std::max(dbl2int((v - lo) / step), 0), dim - 1);
}

}  // namespace

bool hier_grids(const double mbb[4], int nv, int vpr, int nlow, RasterGrid &fine, std::vector<RasterGrid> &coarse) {
    coarse.clear();
    fine = size_grid(mbb, nv / vpr);
    if (!align_grid(fine)) return false;
    const RasterGrid &f = fine;
    // Ideal::layering (Ideal.cpp:1061-1099): coarser layers by doubling the
    // step until the layer has at most nlow pixels
    const int high_layers = static_cast<int>(
        std::ceil(std::max(std::log(f.dimx) / std::log(2.0), std::log(f.dimy) / std::log(2.0))));
    int low_layers = high_layers;
    double dx = f.dimx, dy = f.dimy;
    for (int i = high_layers - 1; i >= 0; i--) {
        if (dx * dy <= nlow) break;
        dx = static_cast<int>(std::ceil(dx / 2.0));
        dy = static_cast<int>(std::ceil(dy / 2.0));
        low_layers--;
    }
    const int nc = high_layers - low_layers;
    coarse.resize(nc);
    for (int i = nc - 1; i >= 0; i--) {
        RasterGrid &g = coarse[i];
        const RasterGrid &finer = i + 1 < nc ? coarse[i + 1] : f;
        const double s0 = finer.step[0] * 2, s1 = finer.step[1] * 2;
        g.step[0] = s0;
        g.step[1] = s1;
        g.low[0] = std::floor(f.low[0] / s0) * s0;
        g.low[1] = std::floor(f.low[1] / s1) * s1;
        g.high[0] = std::ceil(f.high[0] / s0) * s0;
        g.high[1] = std::ceil(f.high[1] / s1) * s1;
        g.dimx = static_cast<int>(std::round((g.high[0] - g.low[0]) / s0));
        g.dimy = static_cast<int>(std::round((g.high[1] - g.low[1]) / s1));
    }
    return true;
}

void build_raster_hier(const float *ring, int nv, const double mbb[4], int vpr, int nlow, RasterOut &out,
                       std::vector<LayerOut> &layers, RasterScratch *S) {
    layers.clear();
    std::vector<RasterGrid> coarse;
    if (!hier_grids(mbb, nv, vpr, nlow, out.g, coarse)) {
        out.ok = false;
        make_empty(out);
        layers.push_back({out.g, {}});
        return;
    }
    build_raster_impl(ring, nv, out, S);
    if (!out.ok) make_empty(out);
    const RasterGrid &f = out.g;
    const int nl = (int)coarse.size() + 1;
    layers.resize(nl);
    layers[nl - 1].g = f;
    for (int i = nl - 2; i >= 0; i--) {
        RasterGrid &g = layers[i].g;
        g = coarse[i];
        const double s0 = g.step[0], s1 = g.step[1];
        // Ideal::merge_status (Ideal.cpp:1040-1059): each coarse pixel sums the
        // finest layer's clipped areas over the fine pixels it covers; it is
        // BORDER if any of them is, else it takes the first one's class
        std::vector<uint8_t> &st = layers[i].status;
        st.assign((size_t)g.dimx * g.dimy, 0);
        const double carea = s0 * s1;
        for (int x = 0; x < g.dimx; x++) {
            for (int y = 0; y < g.dimy; y++) {
                const double lx = g.low[0] + x * s0, ly = g.low[1] + y * s1;
                const double hx = g.low[0] + (x + 1) * s0, hy = g.low[1] + (y + 1) * s1;
                const int x0 = grid_off(lx, f.low[0], f.step[0], f.dimx);
                const int y0 = grid_off(ly, f.low[1], f.step[1], f.dimy);
                int x1 = grid_off(hx, f.low[0], f.step[0], f.dimx) - 1;
                int y1 = grid_off(hy, f.low[1], f.step[1], f.dimy) - 1;
                if (y1 < y0) y1 = y0;
                if (x1 < x0) x1 = x0;
                double area = 0.0;
                bool border = false;
                for (int a = x0; a <= x1; a++)
                    for (int b = y0; b <= y1; b++) {
                        const uint8_t c = out.status[b * f.dimx + a];
                        if (c != 0 && c != kCC - 1) border = true;
                        area += out.areas[b * f.dimx + a];
                    }
                uint8_t cls;
                if (border) {
                    // classifyPixel(area, pixelArea, 16) (geometry_computation.h:299-321)
                    const double ratio = area / carea;
                    if (std::fabs(ratio - 1.0) < 1e-9) cls = kCC - 1;
                    else if (std::fabs(ratio) < 1e-9) cls = 0;
                    else {
                        int idx = static_cast<int>(std::ceil(ratio * (kCC - 2)));
                        if (idx >= kCC) idx = kCC - 1;
                        cls = (uint8_t)idx;
                    }
                } else {
                    const uint8_t c = out.status[y0 * f.dimx + x0];
                    cls = (c == kCC - 1) ? kCC - 1 : 0;
                }
                st[(size_t)y * g.dimx + x] = cls;
            }
        }
    }
}

static void make_empty(RasterOut &out) {
    out.ok = false;
    // The reference raster builder asserts or never terminates on this ring.
    // Keep the grid but make the polygon empty (every pixel OUT, no edge
    // sequences) so that no query reads a half-built raster; the failure is
    // reported through farm_hostset_raster_failures.
    const int npix = out.g.dimx * out.g.dimy;
    out.status.assign(npix, 0);
    out.areas.assign(npix, 0.0);
    out.pixoff.assign((size_t)npix + 1, 0);
    out.seqs.clear();
    out.voff.assign((size_t)out.g.dimx + 2, 0);
    out.vnodes.clear();
    out.hoff.assign((size_t)out.g.dimy + 2, 0);
    out.hnodes.clear();
}

}  // namespace farm
