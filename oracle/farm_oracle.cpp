/*
 * farm_oracle.cpp -- CPU restatement of the reference FARM CPU path
 * (intersect_cpu / within_polygon_cpu and the raster model they build).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg are the only users.
 * It is deliberately written as a literal, unoptimised restatement (std::map,
 * linear dedup scans, one object per polygon), independent of the product's
 * own host raster builder in farm_amd/csrc/host/.
 *
 * Numerics: built with g++ -O2 -std=c++17 -ffp-contract=off, as the
 * reference's Makefile builds its CPU path (src/Makefile:11): IEEE binary32
 * for Point arithmetic, binary64 for boxes, steps and distances, no FMA.
 *
 * How each function is pinned (the reference has no tests and cannot be built
 * here: it needs boost, which is absent; its one data table, the degree-per-km
 * longitude table, is a golden fixture):
 *   double_to_int -> closed forms (tests/test_oracle_primitives.py)
 *   retrieve_pixels (shrunk), the reference GPU filter's unshrunk window,
 *       get_pixel_box -> closed forms on grid lines, 1e-6 and 1e-8 offsets,
 *       clamping; exact rational pixel boxes (tests/test_oracle_windows.py)
 *   segment_intersect -> exact rational orientation on integer grids where
 *       FP32 is exact, plus hand-built collinear/touching cases
 *   .dt decode / MBB -> byte-level format tests (tests/test_host_io.py)
 *   init_raster / rasterization -> closed forms for axis-aligned rectangles
 *       and right triangles (exact fill classes), PIP brute force for IN/OUT
 *       pixel corners, sum(areas) == shoelace area, edge coverage of pixel
 *       edge sequences (tests/test_oracle_raster.py)
 *   grid_align / layering / merge_status -> closed forms on a square snapped
 *       to the universal grid: steps, boxes, dims, layer count and every
 *       merged fill class (tests/test_host_hierarchy.py)
 *   intersect (exact) -> both directions against brute force: (an FP32
 *       crossing of any edge pair) or (an IN pixel overlap in the windows)
 *       (tests/test_oracle_intersect.py)
 *   intersect mode 1 (the reference -a) -> closed forms on rectangles with
 *       known fill classes, and the unshrunk source window
 *       (tests/test_oracle_windows.py)
 *   candidate pairs -> brute-force MBB overlap (independent numpy check)
 *   within / distances -> closed-form point-segment distances, the 900 table
 *       entries (golden fixture), brute-force all-pairs segment distance for
 *       the decision when it is not near the threshold
 *       (tests/test_oracle_within.py, tests/test_oracle_primitives.py)
 */
#include "farm_oracle.h"

#include <algorithm>
#include <atomic>
#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <thread>
#include <tuple>
#include <unordered_set>
#include <vector>

namespace {

/* A reference assert would fire: the reference binary aborts here. */
struct RefAbort : std::runtime_error {
    explicit RefAbort(const char *w) : std::runtime_error(w) {}
};
#define REF_ASSERT(c)                       \
    do {                                    \
        if (!(c)) throw RefAbort(#c);       \
    } while (0)

/* ---------------------------------------------------------------------- */
/* Point (src/include/Point.h:12-97)                                      */
/* ---------------------------------------------------------------------- */
struct Pt {
    float x, y;
};

/* Point::operator== (Point.h:32): float difference, |d| < 1e-9 */
static inline bool pt_eq(const Pt &a, const Pt &b) {
    float dx = a.x - b.x;
    float dy = a.y - b.y;
    return std::fabs(dx) < 1e-9 && std::fabs(dy) < 1e-9;
}
/* Point::operator!= (Point.h:33) */
static inline bool pt_ne(const Pt &a, const Pt &b) {
    float dx = a.x - b.x;
    float dy = a.y - b.y;
    return std::fabs(dx) >= 1e-9 || std::fabs(dy) >= 1e-9;
}

/* Point::cross(a, b) (Point.h:93-97): (a - this) x (b - this), all float */
static inline float cross3(const Pt &p, const Pt &a, const Pt &b) {
    float ax = a.x - p.x, ay = a.y - p.y;
    float bx = b.x - p.x, by = b.y - p.y;
    float l = ax * by;
    float r = ay * bx;
    return l - r;
}

/* sgn (geometry_computation.h:190-193) on a float promoted to double */
static inline int sgn(double x) { return x >= 0 ? (x != 0 ? 1 : 0) : -1; }

/* inter1 (geometry_computation.h:195-212) */
static inline bool inter1(double a, double b, double c, double d) {
    double tmp;
    if (a > b) { tmp = a; a = b; b = tmp; }
    if (c > d) { tmp = c; c = d; d = tmp; }
    return std::max(a, c) <= std::min(b, d);
}

/* segment_intersect (geometry_computation.h:215-221) */
static inline bool segment_intersect(const Pt &a, const Pt &b, const Pt &c, const Pt &d) {
    if (cross3(c, a, d) == 0 && cross3(c, b, d) == 0)
        return inter1(a.x, b.x, c.x, d.x) && inter1(a.y, b.y, c.y, d.y);
    return sgn(cross3(a, b, c)) != sgn(cross3(a, b, d)) && sgn(cross3(c, d, a)) != sgn(cross3(c, d, b));
}

/* segment_intersect_batch (geometry_computation.h:224-237) with a test counter */
static inline bool segment_intersect_batch(const Pt *p1, const Pt *p2, int s1, int s2, int64_t &tests) {
    for (int i = 0; i < s1; i++) {
        for (int j = 0; j < s2; j++) {
            tests++;
            if (segment_intersect(p1[i], p1[i + 1], p2[j], p2[j + 1])) return true;
        }
    }
    return false;
}

/* ---------------------------------------------------------------------- */
/* util.h helpers                                                          */
/* ---------------------------------------------------------------------- */
/* double_to_int (util.h:98-106) */
static inline int double_to_int(double val) {
    int vi = (int)val;
    if (std::abs(1.0 * (vi + 1) - val) < 0.00000001) vi++;
    return vi;
}

/* degree_per_kilometer_latitude / _longitude (util.h:72-96).  The reference
 * embeds a 900-entry table; every entry equals the value of
 * 360 / (sin((90 - k/10) * PI / 180) * 40076) with PI = 3.14159265
 * (util.h:41, 76-85) printed with six decimals, so the table is rebuilt here
 * from that formula (tests pin entries against the published values). */
static const double kDegPerKmLat = 360.0 / 40076.0;
struct DpkTable {
    double v[900];
    DpkTable() {
        const double PI = 3.14159265;
        char buf[64];
        for (int k = 0; k < 900; k++) {
            double f = 360.0 / (std::sin((90 - k / 10.0) * PI / 180) * 40076);
            snprintf(buf, sizeof buf, "%.6f", f);
            v[k] = strtod(buf, nullptr);
        }
    }
};
static const DpkTable kDpk;
static inline double deg_per_km_lon(double latitude) {
    double absla = std::abs(latitude);
    if (absla == 90) absla = 89.9;
    return kDpk.v[(int)(absla * 10)];
}

/* mean[20] (util.h:68): fill-class calibration used by within's ordering */
static const float kMean[20] = {0.0f,  0.7f,  0.62f, 0.55f, 0.52f, 0.45f, 0.42f, 0.4f,  0.38f, 0.34f,
                                0.31f, 0.28f, 0.25f, 0.22f, 0.18f, 0.15f, 0.1f,  0.05f, 0.0f,  0.0f};

/* roundToSignificantDigits (util.h:51-63) */
static double round_sig(double num, int n) {
    if (num == 0) return 0;
    double d = std::ceil(std::log10(num < 0 ? -num : num));
    int power = n - static_cast<int>(d);
    double magnitude = std::pow(10.0, power);
    double shifted = std::round(num * magnitude);
    return shifted / magnitude;
}

/* ---------------------------------------------------------------------- */
/* box (src/include/Box.h, src/geometry/Box.cpp)                           */
/* ---------------------------------------------------------------------- */
struct Box {
    double low[2] = {100000.0, 100000.0};
    double high[2] = {-100000.0, -100000.0};
};

/* box::intersect(box) (Box.cpp:53-58): inclusive */
static inline bool box_intersect(const Box &a, const Box &t) {
    return !(t.low[0] > a.high[0] || t.high[0] < a.low[0] || t.low[1] > a.high[1] || t.high[1] < a.low[1]);
}

/* box::distance(box, geography) (Box.cpp:89-120) */
static double box_distance(const Box &b, const Box &t) {
    if (box_intersect(b, t)) return 0;
    double dx = 0, dy = 0;
    if (t.low[1] > b.high[1]) dy = t.low[1] - b.high[1];
    else if (t.high[1] < b.low[1]) dy = b.low[1] - t.high[1];
    else dy = 0;
    if (t.low[0] > b.high[0]) dx = t.low[0] - b.high[0];
    else if (t.high[0] < b.low[0]) dx = b.low[0] - t.high[0];
    else dx = 0;
    dy = dy / kDegPerKmLat;
    dx = dx / deg_per_km_lon(b.low[1]);
    return std::sqrt(dx * dx + dy * dy);
}

/* box::max_distance(box, geography) (Box.cpp:122-149): corners are float
 * Points, differences are float subtractions promoted to double */
static double box_max_distance(const Box &b, const Box &t) {
    Pt v1[4] = {{(float)b.low[0], (float)b.low[1]},
                {(float)b.low[0], (float)b.high[1]},
                {(float)b.high[0], (float)b.low[1]},
                {(float)b.high[0], (float)b.high[1]}};
    Pt v2[4] = {{(float)t.low[0], (float)t.low[1]},
                {(float)t.low[0], (float)t.high[1]},
                {(float)t.high[0], (float)t.low[1]},
                {(float)t.high[0], (float)t.high[1]}};
    double max_dist = 0.0;
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) {
            float fdx = v1[i].x - v2[j].x;
            float fdy = v1[i].y - v2[j].y;
            double dx = fdx, dy = fdy;
            dy = dy / kDegPerKmLat;
            dx = dx / deg_per_km_lon(b.low[1]);
            max_dist = std::max(max_dist, std::sqrt(dx * dx + dy * dy));
        }
    }
    return max_dist;
}

/* box::expand(buffer, geography) (Box.cpp:200-214) */
static Box box_expand(const Box &in, double buf) {
    Box b = in;
    double shiftx = buf * deg_per_km_lon(b.low[1]);
    double shifty = buf * kDegPerKmLat;
    b.low[0] -= shiftx;
    b.high[0] += shiftx;
    b.low[1] -= shifty;
    b.high[1] += shifty;
    return b;
}

/* point_to_segment_distance (geometry_computation.h:49-89), geography on */
static inline double point_to_segment_distance(const Pt &p, const Pt &p1, const Pt &p2, bool geography) {
    double A = (double)(p.x - p1.x);
    double B = (double)(p.y - p1.y);
    double C = (double)(p2.x - p1.x);
    double D = (double)(p2.y - p1.y);
    double dot = A * C + B * D;
    double len_sq = C * C + D * D;
    double param = -1;
    if (len_sq != 0) param = dot / len_sq;
    double xx, yy;
    if (param < 0) {
        xx = p1.x;
        yy = p1.y;
    } else if (param > 1) {
        xx = p2.x;
        yy = p2.y;
    } else {
        xx = p1.x + param * C;
        yy = p1.y + param * D;
    }
    double dx = p.x - xx;
    double dy = p.y - yy;
    if (geography) {
        dx = dx / deg_per_km_lon(p.y);
        dy = dy / kDegPerKmLat;
    }
    return std::sqrt(dx * dx + dy * dy);
}

/* point_to_segment_within_batch (geometry_computation.h:136-152): note the
 * loop bound seq_len-1 over a length that already counts edges */
static inline double point_to_segment_within_batch(const Pt &p, const Pt *vs, size_t seq_len) {
    double mindist = DBL_MAX;
    for (int i = 0; i < (int)seq_len - 1; i++) {
        double dist = point_to_segment_distance(p, vs[i], vs[i + 1], true);
        if (dist < mindist) mindist = dist;
    }
    return mindist;
}

/* segment_to_segment_within_batch (geometry_computation.h:154-182) */
static inline double segment_to_segment_within_batch(const Pt *vs1, const Pt *vs2, size_t s1, size_t s2,
                                                     int64_t &evals) {
    double mindist = DBL_MAX;
    for (int i = 0; i < (int)s1; i++) {
        double dist = point_to_segment_within_batch(vs1[i], vs2, s2);
        evals += (int64_t)s2 - 1;
        if (dist < mindist) mindist = dist;
    }
    for (int i = 0; i < (int)s2; i++) {
        double dist = point_to_segment_within_batch(vs2[i], vs1, s1);
        evals += (int64_t)s1 - 1;
        if (dist < mindist) mindist = dist;
    }
    return mindist;
}

/* classifyPixel(area, pixelArea, count) (geometry_computation.h:299-321) */
static inline int classify_pixel(double area, double pixelArea, int count) {
    double ratio = area / pixelArea;
    if (std::fabs(ratio - 1.0) < 1e-9) return count - 1;
    if (std::fabs(ratio) < 1e-9) return 0;
    int idx = static_cast<int>(std::ceil(ratio * (count - 2)));
    if (idx >= count) idx = count - 1;
    REF_ASSERT(idx < 256);
    return (uint8_t)idx;
}

/* ---------------------------------------------------------------------- */
/* Raster grid (MyRaster.h/.cpp)                                           */
/* ---------------------------------------------------------------------- */
enum { ST_OUT = 0, ST_BORDER = 1, ST_IN = 2 };
static const int kCC = 16; /* MyRaster::category_count (MyRaster.h:33) */

struct Grid {
    Box mbr;
    int dimx = 0, dimy = 0;
    double step_x = 0, step_y = 0;
    const uint8_t *st = nullptr; /* status bytes */

    /* MyRaster::show_status (MyRaster.cpp:228-239) */
    int show(int id) const {
        if (st[id] == 0) return ST_OUT;
        if (st[id] == kCC - 1) return ST_IN;
        return ST_BORDER;
    }
    /* get_id / get_x / get_y (MyRaster.cpp:170-185) */
    int id(int x, int y) const {
        REF_ASSERT(x >= 0 && x < dimx);
        REF_ASSERT(y >= 0 && y < dimy);
        return y * dimx + x;
    }
    int gx(int id) const { return id % dimx; }
    int gy(int id) const { return id / dimx; }
    /* get_offset_x / _y (MyRaster.cpp:188-200) */
    int off_x(double xval) const {
        int x = double_to_int((xval - mbr.low[0]) / step_x);
        return std::min(std::max(x, 0), dimx - 1);
    }
    int off_y(double yval) const {
        int y = double_to_int((yval - mbr.low[1]) / step_y);
        return std::min(std::max(y, 0), dimy - 1);
    }
    /* get_pixel_box (MyRaster.cpp:312-322) */
    Box pixel_box(int x, int y) const {
        Box b;
        b.low[0] = mbr.low[0] + x * step_x;
        b.low[1] = mbr.low[1] + y * step_y;
        b.high[0] = mbr.low[0] + (x + 1) * step_x;
        b.high[1] = mbr.low[1] + (y + 1) * step_y;
        return b;
    }
    /* retrieve_pixels (MyRaster.cpp:332-346): x-major, then y */
    void retrieve(const Box &t, std::vector<int> &ret) const {
        ret.clear();
        int sx = off_x(t.low[0] + 1e-6);
        int sy = off_y(t.low[1] + 1e-6);
        int ex = off_x(t.high[0] - 1e-6);
        int ey = off_y(t.high[1] - 1e-6);
        for (int i = sx; i <= ex; i++)
            for (int j = sy; j <= ey; j++) ret.push_back(id(i, j));
    }
    /* the reference GPU filter's source window over the target MBB: the same
     * offsets without the 1e-6 shrink (query_intersect.cu:51-54,
     * gpu_get_offset_x/_y geometry.cuh:72-82) */
    void retrieve_unshrunk(const Box &t, std::vector<int> &ret) const {
        ret.clear();
        int sx = off_x(t.low[0]);
        int sy = off_y(t.low[1]);
        int ex = off_x(t.high[0]);
        int ey = off_y(t.high[1]);
        for (int i = sx; i <= ex; i++)
            for (int j = sy; j <= ey; j++) ret.push_back(id(i, j));
    }
    double pixel_area() const { return step_x * step_y; }
};

/* ---------------------------------------------------------------------- */
/* Polygon + IDEAL raster (Ideal.h, Ideal.cpp)                             */
/* ---------------------------------------------------------------------- */
struct CrossInfo {
    int type; /* 0 ENTER, 1 LEAVE */
    int edge_id;
};

struct Layer {
    Grid g;
    std::vector<uint8_t> status; /* own storage (coarse layers) */
};

struct Poly {
    std::vector<Pt> p; /* decoded ring: stored vertices + copy of p[0] */
    Box mbr;           /* getMBB(); grid_align (-h) moves it */
    Box mbb0;          /* the MBB as loaded, which the R-tree indexes */
    Grid g;
    std::vector<uint8_t> status;
    std::vector<double> areas;
    std::vector<uint32_t> offset;  /* dimx*dimy+1; 0xffffffff = never written */
    std::vector<std::pair<uint32_t, uint32_t>> seqs;
    std::vector<uint32_t> hoff, voff;
    std::vector<double> hnodes, vnodes;
    /* hierarchy (-h) */
    bool hier = false;
    bool failed = false; /* the reference raster builder aborts / never ends */
    std::vector<Layer> layers;

    int nv() const { return (int)p.size(); }

    /* Ideal::get_num_sequences (Ideal.cpp:21-26) */
    uint32_t num_seqs(int id) const {
        if (g.show(id) != ST_BORDER) return 0;
        REF_ASSERT(offset[id] != 0xffffffffu && offset[id + 1] != 0xffffffffu);
        return offset[id + 1] - offset[id];
    }
    /* Ideal::decodePixelArea (Ideal.cpp:47-62) */
    double decode_area(int id, bool isLow) const {
        uint8_t f = status[id];
        double pa = g.pixel_area();
        if (f == 0) return 0.0f;
        if (f == kCC - 1) return pa;
        return (1.0 * f - isLow) / (kCC - 2) * pa;
    }
    /* decodePixelArea (Ideal.cpp:47-62) of the fill class the same clipped
     * area gets with cc categories (encodePixelArea, Ideal.cpp:64-88, with cc
     * in place of 16; BORDER classes clamped to [1, cc - 2] as
     * calculate_fullness does, Ideal.cpp:905-912): the reference GPU filter
     * decodes with its category_count (gpu_decode_fullness,
     * geometry.cuh:374-388) */
    double decode_area_cc(int id, bool isLow, int cc) const {
        if (cc == kCC) return decode_area(id, isLow);
        const uint8_t s = status[id];
        const double pa = g.pixel_area();
        if (s == 0) return 0.0f;
        if (s == kCC - 1) return pa;
        const double ratio = areas[id] / pa;
        int f;
        if (std::fabs(ratio - 1.0) < 1e-9) f = cc - 1;
        else if (std::fabs(ratio) < 1e-9) f = 0;
        else {
            f = static_cast<int>(std::ceil(ratio * (cc - 2)));
            if (f >= cc) f = cc - 1;
        }
        f = std::max(1, std::min(f, cc - 2));
        return (1.0 * f - isLow) / (cc - 2) * pa;
    }
};

/* VertexSequence::getMBR (vertex_sequence.cpp:152-167): float bias 1e-9 */
static Box ring_mbr(const std::vector<Pt> &p) {
    Box b;
    float bias = 1e-9f;
    for (size_t i = 0; i < p.size(); i++) {
        b.low[0] = std::min(b.low[0], (double)p[i].x);
        b.high[0] = std::max(b.high[0], (double)p[i].x);
        b.low[1] = std::min(b.low[1], (double)p[i].y);
        b.high[1] = std::max(b.high[1], (double)p[i].y);
    }
    b.low[0] -= bias;
    b.high[0] += bias;
    b.low[1] -= bias;
    b.high[1] += bias;
    return b;
}

/* MyRaster::init_raster(int) (MyRaster.cpp:8-33) */
static void init_raster(Poly &P, int num_pixels) {
    REF_ASSERT(num_pixels >= 0);
    Grid &g = P.g;
    g.mbr = P.mbr;
    double multi = std::abs((P.mbr.high[1] - P.mbr.low[1]) / (P.mbr.high[0] - P.mbr.low[0]));
    g.dimx = (int)std::pow(num_pixels / multi, 0.5);
    g.dimy = (int)(g.dimx * multi);
    if (g.dimx == 0) g.dimx = 1;
    if (g.dimy == 0) g.dimy = 1;
    g.step_x = (P.mbr.high[0] - P.mbr.low[0]) / g.dimx;
    g.step_y = (P.mbr.high[1] - P.mbr.low[1]) / g.dimy;
    if (g.step_x < 0.00001) {
        g.step_x = 0.00001;
        g.dimx = (int)((P.mbr.high[0] - P.mbr.low[0]) / g.step_x + 1);
    }
    if (g.step_y < 0.00001) {
        g.step_y = 0.00001;
        g.dimy = (int)((P.mbr.high[1] - P.mbr.low[1]) / g.step_y + 1);
    }
}

static inline void set_st(Poly &P, int id, uint8_t s) { P.status[id] = s; }

/* Ideal::process_crosses (Ideal.cpp:90-183) */
static void process_crosses(Poly &P, const std::map<int, std::vector<CrossInfo>> &edges_info) {
    int num_edge_seqs = 0;
    for (auto &info : edges_info) {
        std::vector<CrossInfo> crosses = info.second;
        if (crosses.size() == 0) return;
        if (crosses.size() % 2 == 1) crosses.push_back({!crosses.back().type, crosses.back().edge_id});
        int start = 0;
        int end = (int)crosses.size() - 1;
        if (crosses[0].type == 1) {
            REF_ASSERT(crosses[end].type == 0);
            num_edge_seqs += 2;
            start++;
            end--;
        }
        for (int i = start; i <= end; i++) {
            REF_ASSERT(crosses[i].type == 0);
            if (i == end || crosses[i + 1].type == 0) {
                num_edge_seqs++;
            } else {
                num_edge_seqs++;
                i++;
            }
        }
    }
    /* init_edge_sequences (Ideal.cpp:28-33) */
    P.seqs.assign(num_edge_seqs, {0, 0});
    REF_ASSERT(num_edge_seqs < 65536);

    auto add_edge = [&](int idx, int s, int e) {
        REF_ASSERT(e - s + 1 > 0);
        P.seqs[idx] = {(uint32_t)s, (uint32_t)(e - s + 1)};
    };
    int idx = 0;
    for (auto &info : edges_info) {
        int pix = info.first;
        std::vector<CrossInfo> crosses = info.second;
        if (crosses.size() == 0) return;
        if (crosses.size() % 2 == 1) crosses.push_back({!crosses.back().type, crosses.back().edge_id});
        int start = 0;
        int end = (int)crosses.size() - 1;
        P.offset[pix] = idx;
        if (crosses[0].type == 1) {
            add_edge(idx++, 0, crosses[0].edge_id);
            add_edge(idx++, crosses[end].edge_id, P.nv() - 2);
            start++;
            end--;
        }
        for (int i = start; i <= end; i++) {
            if (i == end || crosses[i + 1].type == 0) {
                add_edge(idx++, crosses[i].edge_id, crosses[i].edge_id);
            } else {
                add_edge(idx++, crosses[i].edge_id, crosses[i + 1].edge_id);
                i++;
            }
        }
    }
}

/* Ideal::process_intersection (Ideal.cpp:185-238) */
static void process_intersection(Poly &P, const std::map<int, std::vector<double>> &info, bool horizontal) {
    size_t num_nodes = 0;
    for (auto &i : info) num_nodes += i.second.size();
    std::vector<uint32_t> &off = horizontal ? P.hoff : P.voff;
    std::vector<double> &nodes_out = horizontal ? P.hnodes : P.vnodes;
    nodes_out.assign(num_nodes, 0.0);
    if (!horizontal) off[P.g.dimx + 1] = (uint32_t)num_nodes;
    int idx = 0;
    for (auto &kv : info) {
        std::vector<double> nodes = kv.second;
        std::sort(nodes.begin(), nodes.end());
        off[kv.first] = idx;
        for (double nd : nodes) nodes_out[idx++] = nd;
    }
    if (horizontal) off[P.g.dimy] = idx;
    else off[P.g.dimx] = idx;
}

/* Ideal::evaluate_edges (Ideal.cpp:252-538) */
static void evaluate_edges(Poly &P) {
    std::map<int, std::vector<double>> hinfo, vinfo;
    std::map<int, std::vector<CrossInfo>> einfo;
    const Grid &g = P.g;
    const double start_x = g.mbr.low[0];
    const double start_y = g.mbr.low[1];
    const double step_x = g.step_x, step_y = g.step_y;
    const int dimx = g.dimx, dimy = g.dimy;
    const uint8_t BORDER = ST_BORDER;

    for (int i = 0; i < P.nv() - 1; i++) {
        double x1 = P.p[i].x, y1 = P.p[i].y;
        double x2 = P.p[i + 1].x, y2 = P.p[i + 1].y;
        int cur_startx = double_to_int((x1 - start_x) / step_x);
        int cur_endx = double_to_int((x2 - start_x) / step_x);
        int cur_starty = double_to_int((y1 - start_y) / step_y);
        int cur_endy = double_to_int((y2 - start_y) / step_y);
        if (cur_startx == dimx) cur_startx--;
        if (cur_endx == dimx) cur_endx--;
        if (cur_starty == dimy) cur_starty--;
        if (cur_endy == dimy) cur_endy--;
        REF_ASSERT(!(cur_startx >= dimx || cur_endx >= dimx || cur_starty >= dimy || cur_endy >= dimy));

        set_st(P, g.id(cur_startx, cur_starty), BORDER);
        set_st(P, g.id(cur_endx, cur_endy), BORDER);
        if (cur_startx == cur_endx && cur_starty == cur_endy) continue;

        if (y1 == y2) {
            if (cur_startx < cur_endx) {
                for (int x = cur_startx; x < cur_endx; x++) {
                    vinfo[x + 1].push_back(y1);
                    einfo[g.id(x, cur_starty)].push_back({1, i});
                    einfo[g.id(x + 1, cur_starty)].push_back({0, i});
                    set_st(P, g.id(x, cur_starty), BORDER);
                    set_st(P, g.id(x + 1, cur_starty), BORDER);
                }
            } else {
                for (int x = cur_startx; x > cur_endx; x--) {
                    vinfo[x].push_back(y1);
                    einfo[g.id(x, cur_starty)].push_back({1, i});
                    einfo[g.id(x - 1, cur_starty)].push_back({0, i});
                    set_st(P, g.id(x, cur_starty), BORDER);
                    set_st(P, g.id(x - 1, cur_starty), BORDER);
                }
            }
        } else if (x1 == x2) {
            if (cur_starty < cur_endy) {
                for (int y = cur_starty; y < cur_endy; y++) {
                    hinfo[y + 1].push_back(x1);
                    einfo[g.id(cur_startx, y)].push_back({1, i});
                    einfo[g.id(cur_startx, y + 1)].push_back({0, i});
                    set_st(P, g.id(cur_startx, y), BORDER);
                    set_st(P, g.id(cur_startx, y + 1), BORDER);
                }
            } else {
                for (int y = cur_starty; y > cur_endy; y--) {
                    hinfo[y].push_back(x1);
                    einfo[g.id(cur_startx, y)].push_back({1, i});
                    einfo[g.id(cur_startx, y - 1)].push_back({0, i});
                    set_st(P, g.id(cur_startx, y), BORDER);
                    set_st(P, g.id(cur_startx, y - 1), BORDER);
                }
            }
        } else {
            double a = (y1 - y2) / (x1 - x2);
            double b = (x1 * y2 - x2 * y1) / (x1 - x2);
            int x = cur_startx;
            int y = cur_starty;
            /* a walk that does not terminate would hang the reference; bound
             * it and report the case instead */
            int64_t guard = 4LL * ((int64_t)dimx + dimy) + 16;
            while (x != cur_endx || y != cur_endy) {
                REF_ASSERT(guard-- > 0);
                bool passed = false;
                double yval = 0, xval = 0;
                int cur_x = 0, cur_y = 0;
                if (x != cur_endx) {
                    if (cur_startx < cur_endx) xval = ((double)x + 1) * step_x + start_x;
                    else xval = (double)x * step_x + start_x;
                    yval = xval * a + b;
                    double q = (yval - start_y) / step_y;
                    if (std::abs(q - std::round(q)) < 1e-9) {
                        /* grid-corner branch (Ideal.cpp:397-423) */
                        if (cur_startx < cur_endx && cur_starty < cur_endy) {
                            vinfo[x + 1].push_back(yval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x++, y++;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        } else if (cur_startx < cur_endx && cur_starty > cur_endy) {
                            vinfo[x].push_back(yval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x++, y--;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        } else if (cur_startx > cur_endx && cur_starty < cur_endy) {
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x--, y++;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        } else if (cur_startx > cur_endx && cur_starty > cur_endy) {
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x--, y--;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        }
                        continue;
                    }
                    cur_y = (int)((yval - start_y) / step_y);
                    if (cur_y > std::max(cur_endy, cur_starty)) cur_y = std::max(cur_endy, cur_starty);
                    if (cur_y < std::min(cur_endy, cur_starty)) cur_y = std::min(cur_endy, cur_starty);
                    if (cur_y == y) {
                        passed = true;
                        if (cur_startx < cur_endx) {
                            vinfo[x + 1].push_back(yval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x++;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        } else {
                            vinfo[x].push_back(yval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            x--;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        }
                    }
                }
                if (y != cur_endy) {
                    if (cur_starty < cur_endy) yval = (y + 1) * step_y + start_y;
                    else yval = y * step_y + start_y;
                    xval = (yval - b) / a;
                    cur_x = (int)((xval - start_x) / step_x);
                    if (cur_x > std::max(cur_endx, cur_startx)) cur_x = std::max(cur_endx, cur_startx);
                    if (cur_x < std::min(cur_endx, cur_startx)) cur_x = std::min(cur_endx, cur_startx);
                    if (cur_x == x) {
                        passed = true;
                        if (cur_starty < cur_endy) {
                            hinfo[y + 1].push_back(xval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            y++;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        } else {
                            hinfo[y].push_back(xval);
                            set_st(P, g.id(x, y), BORDER);
                            einfo[g.id(x, y)].push_back({1, i});
                            y--;
                            einfo[g.id(x, y)].push_back({0, i});
                            set_st(P, g.id(x, y), BORDER);
                        }
                    }
                }
                REF_ASSERT(passed);
            }
        }
    }

    if (einfo.size() == 0 && P.nv() > 0) {
        /* special case (Ideal.cpp:527-534): the whole ring in one pixel.
         * The reference stores (0, num_vertices) here, one edge more than
         * the ring has, and its refinement then reads one vertex past the
         * array (undefined).  This restatement keeps the ring's real edges:
         * (0, num_vertices - 1). */
        P.seqs.assign(1, {0u, (uint32_t)(P.nv() - 1)});
        P.offset[0] = 0;
    } else {
        process_crosses(P, einfo);
    }
    process_intersection(P, hinfo, true);
    process_intersection(P, vinfo, false);

    /* process_pixels_null (Ideal.cpp:35-45) with status still 0/1 */
    int n = dimx * dimy;
    P.offset[n] = (uint32_t)P.seqs.size();
    for (int i = n - 1; i >= 0; i--) {
        if (g.show(i) != ST_BORDER) P.offset[i] = P.offset[i + 1];
    }
}

/* Ideal::scanline_reandering (Ideal.cpp:540-573) */
static void scanline(Poly &P) {
    const Grid &g = P.g;
    const double start_x = g.mbr.low[0];
    for (int y = 1; y < g.dimy; y++) {
        bool isin = false;
        uint32_t i = P.hoff[y], j = P.hoff[y + 1];
        for (int x = 0; x < g.dimx; x++) {
            if (g.show(g.id(x, y)) != ST_BORDER) {
                set_st(P, g.id(x, y), isin ? ST_IN : ST_OUT);
                continue;
            }
            int pass = 0;
            while (i < j && P.hnodes[i] <= start_x + g.step_x * (x + 1)) {
                pass++;
                i++;
            }
            if (pass % 2 == 1) isin = !isin;
        }
    }
}

/* TempPolygon::addPoint (MyPolygon.h:49-57): linear tolerant dedup */
static inline void add_point(std::vector<Pt> &v, const Pt &p) {
    for (const Pt &q : v)
        if (pt_eq(q, p)) return;
    v.push_back(p);
}

static inline bool in_v(double Xi, double Xi1, double px) { return !(px < Xi || px > Xi1); }
static inline bool in_h(double Yi, double Yi1, double py) { return !(py < Yi || py > Yi1); }

/* getSlope (geometry_computation.h:23-26): f64 numerator, f32 denominator */
static inline double get_slope(const Pt &A, const Pt &B) {
    float den = B.x - A.x;
    return ((double)B.y - A.y) / den;
}

/* intersectionDualX (Ideal.cpp:575-671) */
static std::vector<Pt> clip_x(const Poly &P, double Xi, double Xi1) {
    std::vector<Pt> out;
    for (int itV = 0; itV < P.nv() - 1; itV++) {
        Pt A = P.p[itV], B = P.p[itV + 1];
        double yi, yi1;
        if (A.x == B.x) {
            yi = std::min(A.y, B.y);
            yi1 = std::min(A.y, B.y);
        } else {
            double slope = get_slope(A, B);
            yi = slope * (Xi - B.x) + B.y;
            yi1 = slope * (Xi1 - B.x) + B.y;
        }
        Pt pi = {(float)Xi, (float)yi};
        Pt pi1 = {(float)Xi1, (float)yi1};
        if (in_v(Xi, Xi1, A.x)) add_point(out, A);
        float mny = std::min(A.y, B.y), mxy = std::max(A.y, B.y);
        float mnx = std::min(A.x, B.x), mxx = std::max(A.x, B.x);
        if (A.x < B.x) {
            if (in_h(mny, mxy, pi.y) && in_v(mnx, mxx, pi.x)) add_point(out, pi);
            if (in_h(mny, mxy, pi1.y) && in_v(mnx, mxx, pi1.x)) add_point(out, pi1);
        } else {
            if (in_h(mny, mxy, pi1.y) && in_v(mnx, mxx, pi1.x)) add_point(out, pi1);
            if (in_h(mny, mxy, pi.y) && in_v(mnx, mxx, pi.x)) add_point(out, pi);
        }
        if (in_v(Xi, Xi1, B.x)) add_point(out, B);
    }
    if (!out.empty()) out.push_back(out.front());
    return out;
}

/* intersectionDualY (Ideal.cpp:673-818) */
static void clip_y(const std::vector<Pt> &pol, double Yi, double Yi1, std::vector<Pt> &out) {
    out.clear();
    for (size_t k = 0; k + 1 < pol.size(); k++) {
        Pt A = pol[k], B = pol[k + 1];
        bool parallels = false;
        double xi = 0, xi1 = 0;
        bool xi_ok = true, xi1_ok = true; /* false stands for -DBL_MAX */
        if (A.x == B.x) {
            xi = A.x;
            xi1 = A.x;
        } else {
            double slope = get_slope(A, B);
            if (A.y == B.y) {
                if (A.y == Yi) {
                    xi = A.x;
                    xi1_ok = false;
                } else if (A.y == Yi1) {
                    xi_ok = false;
                    xi1 = A.x;
                } else {
                    parallels = true;
                }
            } else {
                xi = ((Yi - B.y) / slope) + B.x;
                xi1 = ((Yi1 - B.y) / slope) + B.x;
            }
        }
        if (in_h(Yi, Yi1, A.y)) out.push_back(A);
        if (!parallels) {
            Pt pi = {(float)xi, (float)Yi};
            Pt pi1 = {(float)xi1, (float)Yi1};
            float mny = std::min(A.y, B.y), mxy = std::max(A.y, B.y);
            float mnx = std::min(A.x, B.x), mxx = std::max(A.x, B.x);
            /* -DBL_MAX converts to -inf and never lies inside [mnx, mxx] */
            bool ok_pi = xi_ok && in_v(mnx, mxx, pi.x) && in_h(mny, mxy, pi.y);
            bool ok_pi1 = xi1_ok && in_v(mnx, mxx, pi1.x) && in_h(mny, mxy, pi1.y);
            if (A.y < B.y) {
                if (ok_pi) out.push_back(pi);
                if (ok_pi1) out.push_back(pi1);
            } else {
                if (ok_pi1) out.push_back(pi1);
                if (ok_pi) out.push_back(pi);
            }
        }
    }
    if (!out.empty() && pt_ne(out.front(), out.back())) out.push_back(out.front());
}

/* computePolygonArea (geometry_computation.h:275-297) */
static double polygon_area(const std::vector<Pt> &poly) {
    int n = (int)poly.size();
    if (n < 3) return 0.0;
    double area = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        const Pt &p1 = poly[i];
        const Pt &p2 = poly[i + 1];
        area += ((double)p1.x * p2.y - (double)p2.x * p1.y);
    }
    return std::abs(area) / 2.0;
}

/* Ideal::encodePixelArea (Ideal.cpp:64-88) */
static int encode_area(const Grid &g, double area) {
    double ratio = area / g.pixel_area();
    if (std::fabs(ratio - 1.0) < 1e-9) return kCC - 1;
    if (std::fabs(ratio) < 1e-9) return 0;
    int idx = static_cast<int>(std::ceil(ratio * (kCC - 2)));
    if (idx >= kCC) idx = kCC - 1;
    REF_ASSERT(idx < 256);
    return (uint8_t)idx;
}

/* Ideal::calculate_fullness (Ideal.cpp:820-918) */
static void calculate_fullness(Poly &P) {
    const Grid &g = P.g;
    double Xi = g.mbr.low[0];
    double Xi1 = Xi + g.step_x;
    double kx = Xi + g.dimx * g.step_x;
    std::vector<std::pair<int, std::vector<Pt>>> strips;
    int x = 0;
    while (Xi1 < kx + 1e-9) {
        std::vector<Pt> t = clip_x(P, Xi, Xi1);
        if (t.size() > 0) strips.emplace_back(x, std::move(t));
        Xi += g.step_x;
        Xi1 = Xi + g.step_x;
        x++;
    }
    double ky = g.mbr.low[1] + g.dimy * g.step_y;
    std::vector<Pt> clipped;
    for (auto &s : strips) {
        double Yi = g.mbr.low[1];
        double Yi1 = Yi + g.step_y;
        int y = 0;
        while (Yi1 < ky + 1e-9) {
            clip_y(s.second, Yi, Yi1, clipped);
            double area = polygon_area(clipped);
            int type = encode_area(g, area);
            int pix = g.id(s.first, y);
            if (P.status[pix] == ST_BORDER) {
                P.status[pix] = (uint8_t)std::max(1, std::min(type, kCC - 2));
                P.areas[pix] = area;
            } else if (P.status[pix] == ST_IN) {
                P.status[pix] = kCC - 1;
                P.areas[pix] = g.pixel_area();
            } else {
                P.status[pix] = 0;
                P.areas[pix] = 0.0;
            }
            Yi += g.step_y;
            Yi1 = Yi + g.step_y;
            y++;
        }
    }
}

/* UniversalGrid::configure(25) (UniversalGrid.h:29-35), max_layers 25
 * (query_context.h:218) */
static double universal_step() {
    double s = 360.0 / std::pow(2, 25);
    return round_sig(s, 2);
}

/* MyRaster::grid_align (MyRaster.cpp:444-491): mutates MBB, dims, steps */
static void grid_align(Poly &P) {
    Grid &g = P.g;
    bool flag1 = false, flag2 = false;
    double x = universal_step(), y = x;
    if (g.step_x <= x) g.step_x = x, flag1 = true;
    if (g.step_y <= y) g.step_y = y, flag2 = true;
    for (int i = 25; i > 0; i--) {
        if (!flag1 && g.step_x > x && g.step_x < x * 2) {
            flag1 = true;
            g.step_x = std::abs(g.step_x - x) < std::abs(g.step_x - x * 2) ? x : x * 2;
        }
        if (!flag2 && g.step_y > y && g.step_y < y * 2) {
            flag2 = true;
            g.step_y = std::abs(g.step_y - y) < std::abs(g.step_y - y * 2) ? y : y * 2;
        }
        if (flag1 && flag2) break;
        x *= 2;
        y *= 2;
    }
    REF_ASSERT(flag1 && flag2);
    g.step_x = std::min(g.step_x, g.step_y);
    g.step_y = g.step_x;
    Box &m = P.mbr;
    m.low[0] = std::floor(m.low[0] / g.step_x) * g.step_x;
    m.low[1] = std::floor(m.low[1] / g.step_y) * g.step_y;
    m.high[0] = std::ceil(m.high[0] / g.step_x) * g.step_x;
    m.high[1] = std::ceil(m.high[1] / g.step_y) * g.step_y;
    g.mbr = m;
    g.dimx = static_cast<int>(std::round((m.high[0] - m.low[0]) / g.step_x));
    g.dimy = static_cast<int>(std::round((m.high[1] - m.low[1]) / g.step_y));
}

/* Ideal::layering (Ideal.cpp:1061-1099) + Hraster::init (Hraster.cpp:3-25) */
static void layering(Poly &P, int NLow) {
    Grid &g = P.g;
    int high_layers = static_cast<int>(std::ceil(std::max(std::log(g.dimx) / std::log(2.0), std::log(g.dimy) / std::log(2.0))));
    int low_layers = high_layers;
    double _dimx = g.dimx, _dimy = g.dimy;
    for (int i = high_layers - 1; i >= 0; i--) {
        if (_dimx * _dimy <= NLow) break;
        _dimx = static_cast<int>(std::ceil(_dimx / 2.0));
        _dimy = static_cast<int>(std::ceil(_dimy / 2.0));
        low_layers--;
    }
    int num_layers = high_layers - low_layers + 1;
    P.layers.assign(num_layers, Layer());
    P.layers[num_layers - 1].g = g; /* shares the polygon MBB and status */
    for (int i = num_layers - 2; i >= 0; i--) {
        Layer &L = P.layers[i];
        const Grid &f = P.layers[i + 1].g;
        L.g.step_x = f.step_x * 2;
        L.g.step_y = f.step_y * 2;
        Box m = P.mbr;
        m.low[0] = std::floor(m.low[0] / L.g.step_x) * L.g.step_x;
        m.low[1] = std::floor(m.low[1] / L.g.step_y) * L.g.step_y;
        m.high[0] = std::ceil(m.high[0] / L.g.step_x) * L.g.step_x;
        m.high[1] = std::ceil(m.high[1] / L.g.step_y) * L.g.step_y;
        L.g.mbr = m;
        L.g.dimx = static_cast<int>(std::round((m.high[0] - m.low[0]) / L.g.step_x));
        L.g.dimy = static_cast<int>(std::round((m.high[1] - m.low[1]) / L.g.step_y));
        L.status.assign((size_t)L.g.dimx * L.g.dimy, 0);
    }
}

/* Ideal::merge_area (Ideal.cpp:1008-1038).  The reference leaves the status
 * argument uninitialised when no BORDER pixel is met; this restatement takes
 * the evident intent: BORDER if any merged pixel is BORDER, else the status
 * of the first merged pixel. */
static double merge_area(const Poly &P, const Box &t, int &st) {
    const Grid &g = P.g;
    int sx = g.off_x(t.low[0]);
    int sy = g.off_y(t.low[1]);
    int ex = g.off_x(t.high[0]) - 1;
    int ey = g.off_y(t.high[1]) - 1;
    REF_ASSERT(sx >= 0 && sy >= 0 && ex < g.dimx && ey < g.dimy);
    if (ey < sy) ey = sy;
    if (ex < sx) ex = sx;
    double area = 0.0;
    bool any_border = false;
    for (int i = sx; i <= ex; i++)
        for (int j = sy; j <= ey; j++) {
            int id = g.id(i, j);
            if (g.show(id) == ST_BORDER) any_border = true;
            area += P.areas[id];
        }
    st = any_border ? ST_BORDER : g.show(g.id(sx, sy));
    return area;
}

/* Ideal::merge_status (Ideal.cpp:1040-1059) */
static void merge_status(Poly &P, Layer &L) {
    L.g.st = L.status.data();
    for (int x = 0; x < L.g.dimx; x++) {
        for (int y = 0; y < L.g.dimy; y++) {
            int st;
            double merged = merge_area(P, L.g.pixel_box(x, y), st);
            uint8_t f;
            if (st == ST_BORDER) f = (uint8_t)classify_pixel(merged, L.g.step_x * L.g.step_y, kCC);
            else if (st == ST_IN) f = kCC - 1;
            else f = 0;
            L.status[L.g.id(x, y)] = f;
        }
    }
}

/* preprocessing.cpp:4-29 rasterization_unit + Ideal::rasterization
 * (Ideal.cpp:920-955) */
static void rasterize(Poly &P, int vpr, bool hier, int nlow) {
    init_raster(P, P.nv() / vpr);
    P.hier = hier;
    if (hier) {
        grid_align(P);
        layering(P, nlow);
    }
    Grid &g = P.g;
    int n = g.dimx * g.dimy;
    P.status.assign(n, 0);
    P.areas.assign(n, 0.0);
    P.offset.assign((size_t)n + 1, 0xffffffffu);
    P.hoff.assign((size_t)g.dimy + 2, 0);
    P.voff.assign((size_t)g.dimx + 2, 0);
    g.st = P.status.data();
    evaluate_edges(P);
    scanline(P);
    calculate_fullness(P);
    if (hier) {
        Layer &last = P.layers.back();
        last.g = g;
        last.g.st = P.status.data();
        for (int i = (int)P.layers.size() - 2; i >= 0; i--) merge_status(P, P.layers[i]);
    }
}

/* ---------------------------------------------------------------------- */
/* Queries                                                                 */
/* ---------------------------------------------------------------------- */
struct QStats {
    int64_t bb_pairs = 0, seg_tests = 0, in_hits = 0;
};

/* Ideal::intersect (Ideal.cpp:1329-1422); mode 1 = the reference's -a, which
 * only its GPU path has: kernel_filter_intersect's windows (the source window
 * over the target MBB unshrunk, query_intersect.cu:51-54; the target window
 * over a source pixel shrunk by 1e-6, :66-74), then true on an IN pixel on
 * either side or on a BORDER x BORDER pair with probability > 1
 * (query_intersect.cu:86-88,124-129) */
static bool intersect(const Poly &A, const Poly &B, int mode, QStats &qs, int cc = kCC) {
    std::vector<std::tuple<double, int, int>> cand;
    std::vector<int> pxs, tpxs;
    if (mode == 1)
        A.g.retrieve_unshrunk(B.mbr, pxs);
    else
        A.g.retrieve(B.mbr, pxs);
    for (int pa : pxs) {
        if (A.g.show(pa) == ST_OUT) continue;
        Box bx = A.g.pixel_box(A.g.gx(pa), A.g.gy(pa));
        B.g.retrieve(bx, tpxs);
        for (int pb : tpxs) {
            if (B.g.show(pb) == ST_OUT) continue;
            if (A.g.show(pa) == ST_IN || B.g.show(pb) == ST_IN) {
                qs.in_hits++;
                return true;
            }
            double s_apx = (A.decode_area_cc(pa, true, cc) + A.decode_area_cc(pa, false, cc)) / 2;
            double t_apx = (B.decode_area_cc(pb, true, cc) + B.decode_area_cc(pb, false, cc)) / 2;
            double prob = (s_apx + t_apx) / std::max(A.g.pixel_area(), B.g.pixel_area());
            cand.emplace_back(prob, pa, pb);
        }
    }
    qs.bb_pairs += (int64_t)cand.size();
    if (mode == 1) {
        for (auto &c : cand)
            if (std::get<0>(c) > 1.0) return true;
        return false;
    }
    std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a > b; });
    for (auto &c : cand) {
        int pa = std::get<1>(c), pb = std::get<2>(c);
        uint32_t na = A.num_seqs(pa), nb = B.num_seqs(pb);
        for (uint32_t i = 0; i < na; i++) {
            auto r = A.seqs[A.offset[pa] + i];
            for (uint32_t j = 0; j < nb; j++) {
                auto r2 = B.seqs[B.offset[pb] + j];
                if (segment_intersect_batch(A.p.data() + r.first, B.p.data() + r2.first, (int)r.second,
                                            (int)r2.second, qs.seg_tests))
                    return true;
            }
        }
    }
    return false;
}

/* Ideal::within (Ideal.cpp:2475-2643): hierarchical BFS over layer pairs,
 * candidate pixel pairs ordered by dist_apx, truncated segment distances.
 * `min_dist_out` receives the running minimum distance when refinement ran
 * (100000 otherwise).  Requires rasterize(..., hierarchy = true). */
static bool within(const Poly &S, const Poly &T, double d, QStats &qs, double *min_dist_out) {
    *min_dist_out = 100000.0;
    REF_ASSERT(S.hier && T.hier && !S.layers.empty() && !T.layers.empty());
    const int s_level = (int)S.layers.size() - 1, t_level = (int)T.layers.size() - 1;
    std::queue<std::pair<int, int>> pairs;
    const Grid &s0 = S.layers[0].g, &t0 = T.layers[0].g;
    for (int a = 0; a < s0.dimx * s0.dimy; a++)
        for (int b = 0; b < t0.dimx * t0.dimy; b++) pairs.push({a, b});
    int i = 0, j = 0;
    while (true) {
        std::unordered_set<int> s_pxs, t_pxs;
        bool s_next = false, t_next = false;
        const double s_step = S.layers[i].g.step_x, t_step = T.layers[j].g.step_x;
        if (i < s_level && (s_step >= t_step || j >= t_level)) {
            i++;
            s_next = true;
        }
        if (j < t_level && (s_step <= t_step || i >= s_level)) {
            j++;
            t_next = true;
        }
        const int size = (int)pairs.size();
        if (size == 0) break;
        const Grid &sg = S.layers[i].g, &tg = T.layers[j].g;
        std::vector<int> tmp;
        for (int k = 0; k < size; k++) {
            auto pr = pairs.front();
            pairs.pop();
            if (s_next) {
                const Grid &pg = S.layers[i - 1].g;
                Box b = pg.pixel_box(pg.gx(pr.first), pg.gy(pr.first));
                b.low[0] += 1e-6;
                b.low[1] += 1e-6;
                b.high[0] -= 1e-6;
                b.high[1] -= 1e-6;
                sg.retrieve(b, tmp);
                for (int p : tmp)
                    if (sg.show(p) == ST_BORDER) s_pxs.insert(p);
            } else if (sg.show(pr.first) == ST_BORDER) {
                s_pxs.insert(pr.first);
            }
            if (t_next) {
                const Grid &pg = T.layers[j - 1].g;
                Box b = pg.pixel_box(pg.gx(pr.second), pg.gy(pr.second));
                b.low[0] += 1e-6;
                b.low[1] += 1e-6;
                b.high[0] -= 1e-6;
                b.high[1] -= 1e-6;
                tg.retrieve(b, tmp);
                for (int p : tmp)
                    if (tg.show(p) == ST_BORDER) t_pxs.insert(p);
            } else if (tg.show(pr.second) == ST_BORDER) {
                t_pxs.insert(pr.second);
            }
        }
        double max_box_dist = 100000.0;
        for (int id1 : s_pxs) {
            Box b1 = sg.pixel_box(sg.gx(id1), sg.gy(id1));
            for (int id2 : t_pxs) {
                Box b2 = tg.pixel_box(tg.gx(id2), tg.gy(id2));
                double mx = box_max_distance(b1, b2);
                if (mx <= d) return true;
                max_box_dist = std::min(max_box_dist, mx);
            }
        }
        for (int id1 : s_pxs) {
            Box b1 = sg.pixel_box(sg.gx(id1), sg.gy(id1));
            for (int id2 : t_pxs) {
                Box b2 = tg.pixel_box(tg.gx(id2), tg.gy(id2));
                double mn = box_distance(b1, b2);
                if (mn > d) continue;
                if (mn < max_box_dist) pairs.push({id1, id2});
            }
        }
        if (!s_next && !t_next) break;
    }
    std::vector<std::tuple<double, double, double, int, int>> cand;
    const double sa = S.g.pixel_area(), ta = T.g.pixel_area();
    while (!pairs.empty()) {
        auto pr = pairs.front();
        pairs.pop();
        const int id1 = pr.first, id2 = pr.second;
        const int pf = classify_pixel(S.areas[id1] + T.areas[id2], sa + ta, 20);
        Box b1 = S.g.pixel_box(S.g.gx(id1), S.g.gy(id1));
        Box b2 = T.g.pixel_box(T.g.gx(id2), T.g.gy(id2));
        const double lo = box_distance(b1, b2);
        const double hi = box_max_distance(b1, b2);
        const double apx = lo + kMean[pf] * (hi - lo);
        cand.emplace_back(apx, lo, hi, id1, id2);
    }
    std::sort(cand.begin(), cand.end());
    qs.bb_pairs += (int64_t)cand.size();
    double min_dist = 100000.0;
    for (size_t pid = 0; pid < cand.size(); pid++) {
        double reference_dist = 100000;
        for (size_t k = pid + 1; k < cand.size(); k++) reference_dist = std::min(reference_dist, std::get<1>(cand[k]));
        const int id1 = std::get<3>(cand[pid]), id2 = std::get<4>(cand[pid]);
        const uint32_t n1 = S.num_seqs(id1), n2 = T.num_seqs(id2);
        for (uint32_t a = 0; a < n1; a++) {
            auto er1 = S.seqs[S.offset[id1] + a];
            for (uint32_t b = 0; b < n2; b++) {
                auto er2 = T.seqs[T.offset[id2] + b];
                if (er1.second < 2 || er2.second < 2) continue;
                double dist = segment_to_segment_within_batch(T.p.data() + er2.first, S.p.data() + er1.first,
                                                              er2.second, er1.second, qs.seg_tests);
                min_dist = std::min(dist, min_dist);
                *min_dist_out = min_dist;
                if (min_dist < d) return true;
                if (min_dist < reference_dist) return false;
            }
        }
    }
    return false;
}

/* ---------------------------------------------------------------------- */
/* Intersection                                                            */
/* ---------------------------------------------------------------------- */
/* Intersection record (Point.h:132-139) */
struct IxRec {
    Pt p;
    int ei, ej;
    double t, u;
};

/* segment_intersect_batch(..., inters) (geometry_computation.h:239-267):
 * f32 edge vectors and crosses, f64 parameters, f32 intersection point */
static void segment_intersect_collect(const Pt *p1, const Pt *p2, int s1, int s2, int e1, int e2,
                                      std::vector<IxRec> &out) {
    for (int i = s1; i < e1; i++) {
        for (int j = s2; j < e2; j++) {
            const float d1x = p1[i + 1].x - p1[i].x, d1y = p1[i + 1].y - p1[i].y;
            const float d2x = p2[j + 1].x - p2[j].x, d2y = p2[j + 1].y - p2[j].y;
            const float rx = p2[j].x - p1[i].x, ry = p2[j].y - p1[i].y;
            const float dn = d1x * d2y - d1y * d2x;
            const double denom = dn;
            if (std::abs(denom) < 1e-12) continue;
            const float c2 = rx * d2y - ry * d2x;
            const float c1 = rx * d1y - ry * d1x;
            const double t = c2 / denom;
            const double u = c1 / denom;
            if (t >= -1e-9 && t <= 1 + 1e-9 && u >= -1e-9 && u <= 1 + 1e-9) {
                const float tf = (float)t; /* Point::operator*(const float &) */
                IxRec r;
                r.p.x = p1[i].x + tf * d1x;
                r.p.y = p1[i].y + tf * d1y;
                r.ei = i;
                r.ej = j;
                r.t = t;
                r.u = u;
                out.push_back(r);
            }
        }
    }
}

/* Ideal::segment_contain (Ideal.cpp:1233-1327) with count_intersection_nodes
 * (Ideal.cpp:988-1006): 0 OUT, 1 BORDER (on the boundary), 2 IN */
static int segment_contain(const Poly &P, const Pt &p) {
    const Grid &g = P.g;
    const int px = g.off_x(p.x), py = g.off_y(p.y);
    const int id = g.id(px, py);
    const Box bx = g.pixel_box(px, py);
    const int st = g.show(id);
    if (st == ST_IN) return 2;
    if (st == ST_OUT) return 0;
    bool ret = false;
    const uint32_t ns = P.num_seqs(id);
    for (uint32_t e = 0; e < ns; e++) {
        const auto sq = P.seqs[P.offset[id] + e];
        const uint32_t pos = sq.first;
        for (uint32_t k = 0; k < sq.second; k++) {
            const Pt &v1 = P.p[pos + k];
            const Pt &v2 = P.p[pos + k + 1];
            if (pt_eq(p, v1) || pt_eq(p, v2)) return 1;
            if ((v1.y >= p.y) != (v2.y >= p.y)) {
                const double dx = v2.x - v1.x; /* f32 differences, widened */
                const double dy = v2.y - v1.y;
                const double py_diff = p.y - v1.y;
                if (std::abs(dy) > 1e-9) {
                    const double int_x = dx * py_diff / dy + v1.x;
                    if (std::fabs(p.x - int_x) < 1e-9) return 1;
                    if (p.x < int_x && int_x <= bx.high[0]) ret = !ret;
                }
            } else if (v1.y == p.y && v2.y == p.y && (v1.x >= p.x) != (v2.x >= p.x)) {
                return 1;
            }
        }
    }
    /* crossing nodes of the vertical grid line right of the pixel */
    int count = 0;
    const int x = px + 1;
    uint32_t i = P.voff[x], j = P.voff[x + 1];
    while (i < j && P.vnodes[i] < p.y) {
        count++;
        i++;
    }
    if (count % 2 == 1) ret = !ret;
    return ret ? 2 : 0;
}

/* Cross-product sum of the path a -> ring vertices -> b relative to origin o,
 * and its representative point: the midpoint (f32) of the longest
 * sub-segment of the path (squared length in f64 from the f32 points, the
 * first one on ties), a point well inside the piece.  first/last: ring
 * vertex range (cyclic over nv - 1 vertices), nvert vertices in it. */
static double path_cross(const Poly &P, const Pt &a, const Pt &b, int first, int nvert, double ox, double oy,
                         Pt *rep) {
    const int m = P.nv() - 1;
    auto at = [&](int k) -> Pt { /* path point k: 0 = a, 1..nvert = vertices, nvert + 1 = b */
        if (k == 0) return a;
        if (k == nvert + 1) return b;
        return P.p[(first + k - 1) % m];
    };
    double sum = 0.0;
    for (int k = 0; k <= nvert; k++) {
        const Pt u = at(k), v = at(k + 1);
        sum += ((double)u.x - ox) * ((double)v.y - oy) - ((double)v.x - ox) * ((double)u.y - oy);
    }
    int h = 0;
    double best = -1.0;
    for (int k = 0; k <= nvert; k++) {
        const Pt u = at(k), v = at(k + 1);
        const double dx = (double)v.x - u.x, dy = (double)v.y - u.y;
        const double l = dx * dx + dy * dy;
        if (l > best) {
            best = l;
            h = k;
        }
    }
    const Pt u = at(h), v = at(h + 1);
    rep->x = (u.x + v.x) * 0.5f;
    rep->y = (u.y + v.y) * 0.5f;
    return sum;
}

/* Shoelace sum of the whole ring relative to its own MBB's low corner
 * (signed, twice the area) */
static double ring_cross(const Poly &P) {
    const double ox = P.mbb0.low[0], oy = P.mbb0.low[1];
    double sum = 0.0;
    for (int k = 0; k + 1 < P.nv(); k++)
        sum += ((double)P.p[k].x - ox) * ((double)P.p[k + 1].y - oy) -
               ((double)P.p[k + 1].x - ox) * ((double)P.p[k].y - oy);
    return sum;
}

struct IxOut {
    int32_t n_inters = 0; /* -1: every source pixel in the window OUT, -2: every one IN */
    int32_t n_segments = 0, n_in = 0, n_out = 0, n_border = 0;
    double area = 0.0;
};

/* Ideal::intersection (Ideal.cpp:1440-1683) up to the segment statuses (the
 * reference returns there, before any polygon assembly), plus the area of
 * A n B by Green's theorem over the classified boundary pieces:
 *   area = |sA * sum(A pieces inside B) + sB * sum(B pieces inside A)| / 2
 * where a piece runs along one ring between consecutive intersection points
 * (ordered by edge, then parameter), its sum is the cross-product sum of its
 * path relative to the low corner of the two MBBs' overlap, it counts when
 * the midpoint of its longest sub-segment is IN the other polygon
 * (segment_contain; BORDER pieces are not credited), and sX = sign of ring
 * X's signed area.  With fewer than two distinct intersection points the
 * area is area(A) if A's first vertex is IN B, area(B) if B's is IN A, else
 * 0; a window of only IN source pixels gives area(B), only OUT 0.  Whole
 * ring areas are shoelace sums relative to the ring's own MBB low corner. */
static IxOut intersection(const Poly &A, const Poly &B) {
    IxOut o;
    const double ox = std::max(A.mbr.low[0], B.mbr.low[0]);
    const double oy = std::max(A.mbr.low[1], B.mbr.low[1]);
    std::vector<int> pxs, tpxs;
    A.g.retrieve(B.mbr, pxs);
    int etn = 0, itn = 0;
    for (int p : pxs) {
        if (A.g.show(p) == ST_OUT) etn++;
        else if (A.g.show(p) == ST_IN) itn++;
    }
    if (etn == (int)pxs.size()) {
        o.n_inters = -1;
        return o;
    }
    if (itn == (int)pxs.size()) {
        o.n_inters = -2;
        o.area = std::fabs(ring_cross(B)) / 2.0;
        return o;
    }
    std::vector<IxRec> in;
    for (int p : pxs) {
        const Box bx = A.g.pixel_box(A.g.gx(p), A.g.gy(p));
        B.g.retrieve(bx, tpxs);
        for (int p2 : tpxs) {
            if (A.g.show(p) != ST_BORDER || B.g.show(p2) != ST_BORDER) continue;
            for (uint32_t i = 0; i < A.num_seqs(p); i++) {
                const auto r = A.seqs[A.offset[p] + i];
                for (uint32_t j = 0; j < B.num_seqs(p2); j++) {
                    const auto r2 = B.seqs[B.offset[p2] + j];
                    REF_ASSERT(r.second != 0 && r2.second != 0);
                    segment_intersect_collect(A.p.data(), B.p.data(), (int)r.first, (int)r2.first,
                                              (int)(r.first + r.second), (int)(r2.first + r2.second), in);
                }
            }
        }
    }
    /* sort by point, drop consecutive equal points (std::sort + std::unique;
     * records with identical points are ordered by (ei, ej, t, u) here, which
     * the reference leaves unspecified) */
    auto key_pt = [](const IxRec &a, const IxRec &b) {
        if (a.p.x != b.p.x) return a.p.x < b.p.x;
        if (a.p.y != b.p.y) return a.p.y < b.p.y;
        if (a.ei != b.ei) return a.ei < b.ei;
        if (a.ej != b.ej) return a.ej < b.ej;
        if (a.t != b.t) return a.t < b.t;
        return a.u < b.u;
    };
    std::sort(in.begin(), in.end(), key_pt);
    std::vector<IxRec> ix;
    for (const IxRec &r : in)
        if (ix.empty() || !pt_eq(ix.back().p, r.p)) ix.push_back(r);
    const int n = (int)ix.size();
    o.n_inters = n;
    const double eps = (double)1e-6f; /* util.h:44 */
    double side[2] = {0.0, 0.0};
    int pieces = 0;
    for (int s = 0; s < 2; s++) {
        const bool src = s == 0;
        const Poly &P = src ? A : B;
        const Poly &Q = src ? B : A;
        std::sort(ix.begin(), ix.end(), [src](const IxRec &a, const IxRec &b) {
            const int ea = src ? a.ei : a.ej, eb = src ? b.ei : b.ej;
            const double ta = src ? a.t : a.u, tb = src ? b.t : b.u;
            if (ea != eb) return ea < eb;
            if (ta != tb) return ta < tb;
            return src ? (a.ej != b.ej ? a.ej < b.ej : a.u < b.u) : (a.ei != b.ei ? a.ei < b.ei : a.t < b.t);
        });
        const int m = P.nv() - 1;
        for (int i = 0; i < n; i++) {
            const IxRec &a = ix[i], &b = ix[(i + 1) % n];
            if (pt_eq(a.p, b.p)) continue;
            /* the reference's segment (Ideal.cpp:1559-1581 / 1592-1614) */
            int ae = src ? a.ei : a.ej, be = src ? b.ei : b.ej;
            double ap = src ? a.t : a.u, bp = src ? b.t : b.u;
            if (std::fabs(ap - 1.0) < eps) {
                ae = (ae + 1) % m;
                ap = 0.0;
            }
            if (std::fabs(bp) < eps) {
                be--;
                bp = 1.0;
            }
            const bool same = ae == be && ap < bp;
            const int edge_start = same ? -1 : ae + 1;
            Pt rp;
            if (edge_start == -1) {
                rp.x = (a.p.x + b.p.x) * 0.5f;
                rp.y = (a.p.y + b.p.y) * 0.5f;
            } else {
                rp = P.p[edge_start];
            }
            const int st = segment_contain(Q, rp);
            o.n_segments++;
            if (st == 2) o.n_in++;
            else if (st == 0) o.n_out++;
            else o.n_border++;
            /* the area piece: a -> vertices (ea0 + 1 .. eb0) -> b along the ring */
            const int ea0 = src ? a.ei : a.ej, eb0 = src ? b.ei : b.ej;
            const double ta0 = src ? a.t : a.u, tb0 = src ? b.t : b.u;
            int nvert;
            if (i + 1 < n) nvert = eb0 - ea0;                      /* later in ring order */
            else nvert = (eb0 - ea0 + m) % m + ((eb0 == ea0 && tb0 <= ta0) ? m : 0); /* wraps */
            if (nvert < 0) nvert = 0;
            Pt rep;
            const double c = path_cross(P, a.p, b.p, ea0 + 1, nvert, ox, oy, &rep);
            if (segment_contain(Q, rep) == 2) side[s] += c;
            pieces++;
        }
    }
    if (pieces == 0) {
        /* no two distinct intersection points: containment or disjoint */
        if (segment_contain(B, A.p[0]) == 2) o.area = std::fabs(ring_cross(A)) / 2.0;
        else if (segment_contain(A, B.p[0]) == 2) o.area = std::fabs(ring_cross(B)) / 2.0;
        return o;
    }
    const double sa = ring_cross(A) < 0 ? -1.0 : 1.0;
    const double sb = ring_cross(B) < 0 ? -1.0 : 1.0;
    o.area = std::fabs(sa * side[0] + sb * side[1]) / 2.0;
    return o;
}

} // namespace

/* ---------------------------------------------------------------------- */
/* C API                                                                   */
/* ---------------------------------------------------------------------- */
struct orc_set {
    std::vector<Poly> polys;
};

extern "C" {

orc_set *orc_create(int64_t n, const uint64_t *ring_off, const float *xy) {
    orc_set *s = new orc_set();
    s->polys.reserve(n);
    for (int64_t i = 0; i < n; i++) {
        uint64_t b = ring_off[i], e = ring_off[i + 1];
        /* VertexSequence::decode (vertex_sequence.cpp:206-218) appends p[0];
         * load_unit keeps polygons with >= 3 decoded vertices
         * (storage.cpp:341) */
        if (e - b + 1 < 3) continue;
        Poly P;
        P.p.resize(e - b + 1);
        for (uint64_t k = b; k < e; k++) P.p[k - b] = {xy[2 * k], xy[2 * k + 1]};
        P.p[e - b] = P.p[0];
        P.mbr = ring_mbr(P.p);
        P.mbb0 = P.mbr;
        s->polys.push_back(std::move(P));
    }
    return s;
}

void orc_destroy(orc_set *s) { delete s; }
int64_t orc_size(const orc_set *s) { return (int64_t)s->polys.size(); }

int orc_mbb(const orc_set *s, int64_t i, double *m) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return ORC_EINVAL;
    const Box &b = s->polys[i].mbr;
    m[0] = b.low[0];
    m[1] = b.low[1];
    m[2] = b.high[0];
    m[3] = b.high[1];
    return ORC_OK;
}

int orc_rasterize(orc_set *s, int vpr, int hierarchy, int nlow, int nthreads, int64_t *bad) {
    if (vpr <= 0) return ORC_EINVAL;
    std::atomic<int64_t> next(0);
    std::atomic<int64_t> first_bad(INT64_MAX);
    int64_t n = (int64_t)s->polys.size();
    auto work = [&]() {
        for (;;) {
            int64_t b = next.fetch_add(10);
            if (b >= n) break;
            for (int64_t i = b; i < std::min(n, b + 10); i++) {
                try {
                    rasterize(s->polys[i], vpr, hierarchy != 0, nlow);
                } catch (const RefAbort &) {
                    s->polys[i].failed = true;
                    int64_t cur = first_bad.load();
                    while (i < cur && !first_bad.compare_exchange_weak(cur, i)) {
                    }
                }
            }
        }
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    if (first_bad.load() != INT64_MAX) {
        if (bad) *bad = first_bad.load();
        return ORC_EREF_ABORT;
    }
    return ORC_OK;
}

int orc_raster_info(const orc_set *s, int64_t i, double *mbr4, int *dims2, double *steps2, int64_t *counts) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return ORC_EINVAL;
    const Poly &P = s->polys[i];
    mbr4[0] = P.g.mbr.low[0];
    mbr4[1] = P.g.mbr.low[1];
    mbr4[2] = P.g.mbr.high[0];
    mbr4[3] = P.g.mbr.high[1];
    dims2[0] = P.g.dimx;
    dims2[1] = P.g.dimy;
    steps2[0] = P.g.step_x;
    steps2[1] = P.g.step_y;
    counts[0] = (int64_t)P.status.size();
    counts[1] = (int64_t)P.seqs.size();
    counts[2] = (int64_t)P.vnodes.size();
    counts[3] = (int64_t)P.hnodes.size();
    counts[4] = (int64_t)P.p.size();
    return ORC_OK;
}

/* MyRaster::retrieve_pixels (unshrunk = 0) or the reference GPU filter's
 * unshrunk window (unshrunk = 1) of box4 in polygon i's raster: pixel ids in
 * enumeration order; returns their count (at most cap are written), or -1 */
int64_t orc_retrieve(const orc_set *s, int64_t i, const double *box4, int unshrunk, int32_t *ids, int64_t cap) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return -1;
    const Poly &P = s->polys[i];
    Box b;
    b.low[0] = box4[0];
    b.low[1] = box4[1];
    b.high[0] = box4[2];
    b.high[1] = box4[3];
    std::vector<int> ids_v;
    if (unshrunk)
        P.g.retrieve_unshrunk(b, ids_v);
    else
        P.g.retrieve(b, ids_v);
    for (int64_t k = 0; k < (int64_t)ids_v.size() && k < cap; k++) ids[k] = ids_v[k];
    return (int64_t)ids_v.size();
}

/* MyRaster::get_pixel_box of pixel (x, y) of polygon i */
int orc_pixel_box(const orc_set *s, int64_t i, int x, int y, double *out4) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return ORC_EINVAL;
    const Box b = s->polys[i].g.pixel_box(x, y);
    out4[0] = b.low[0];
    out4[1] = b.low[1];
    out4[2] = b.high[0];
    out4[3] = b.high[1];
    return ORC_OK;
}

int orc_raster_copy(const orc_set *s, int64_t i, uint8_t *status, uint32_t *offset, uint32_t *seqs,
                    uint32_t *voff, double *vnodes, uint32_t *hoff, double *hnodes, double *areas) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return ORC_EINVAL;
    const Poly &P = s->polys[i];
    if (status) memcpy(status, P.status.data(), P.status.size());
    if (offset) memcpy(offset, P.offset.data(), P.offset.size() * 4);
    if (seqs)
        for (size_t k = 0; k < P.seqs.size(); k++) {
            seqs[2 * k] = P.seqs[k].first;
            seqs[2 * k + 1] = P.seqs[k].second;
        }
    if (voff) memcpy(voff, P.voff.data(), P.voff.size() * 4);
    if (vnodes) memcpy(vnodes, P.vnodes.data(), P.vnodes.size() * 8);
    if (hoff) memcpy(hoff, P.hoff.data(), P.hoff.size() * 4);
    if (hnodes) memcpy(hnodes, P.hnodes.data(), P.hnodes.size() * 8);
    if (areas) memcpy(areas, P.areas.data(), P.areas.size() * 8);
    return ORC_OK;
}

int orc_layer(const orc_set *s, int64_t i, int k, double *low2, double *step, int *dims2, uint8_t *status) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return -1;
    const Poly &P = s->polys[i];
    const int nl = (int)P.layers.size();
    if (k < 0 || k >= nl) return nl;
    const Grid &g = P.layers[k].g;
    low2[0] = g.mbr.low[0];
    low2[1] = g.mbr.low[1];
    *step = g.step_x;
    dims2[0] = g.dimx;
    dims2[1] = g.dimy;
    if (status) {
        if (g.st) memcpy(status, g.st, (size_t)g.dimx * g.dimy);
        else memset(status, 0, (size_t)g.dimx * g.dimy);
    }
    return nl;
}

int orc_intersect(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa, const uint32_t *pb,
                  uint8_t *out, int mode, int nthreads, int64_t *stats) {
    for (int64_t k = 0; k < npairs; k++)
        if (pa[k] >= a->polys.size() || pb[k] >= b->polys.size()) return ORC_EINVAL;
    /* mode bits 8..15: b of the -w b fill categories (cc = 2^b, 0 = 16) */
    const int wbits = (mode >> 8) & 0xff;
    if (wbits != 0 && (wbits < 2 || wbits > 8)) return ORC_EINVAL;
    const int cc = wbits ? 1 << wbits : kCC;
    std::atomic<int64_t> next(0);
    std::atomic<int> err(ORC_OK);
    std::mutex mu;
    QStats total;
    auto work = [&]() {
        QStats qs;
        for (;;) {
            /* query_context::next_batch(10) (query_context.cpp:145-160) */
            int64_t s = next.fetch_add(10);
            if (s >= npairs) break;
            for (int64_t k = s; k < std::min(npairs, s + 10); k++) {
                try {
                    if (a->polys[pa[k]].failed || b->polys[pb[k]].failed) {
                        out[k] = 2;
                        err = ORC_EREF_ABORT;
                        continue;
                    }
                    out[k] = intersect(a->polys[pa[k]], b->polys[pb[k]], mode & 0xff, qs, cc) ? 1 : 0;
                } catch (const RefAbort &) {
                    out[k] = 2;
                    err = ORC_EREF_ABORT;
                }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        total.bb_pairs += qs.bb_pairs;
        total.seg_tests += qs.seg_tests;
        total.in_hits += qs.in_hits;
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    if (stats) {
        stats[0] = total.bb_pairs;
        stats[1] = total.seg_tests;
        stats[2] = total.in_hits;
    }
    return err.load();
}

int orc_intersection(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa, const uint32_t *pb,
                     int32_t *counts, double *area, int nthreads) {
    for (int64_t k = 0; k < npairs; k++)
        if (pa[k] >= a->polys.size() || pb[k] >= b->polys.size()) return ORC_EINVAL;
    std::atomic<int64_t> next(0);
    std::atomic<int> err(ORC_OK);
    auto work = [&]() {
        for (;;) {
            int64_t s = next.fetch_add(10);
            if (s >= npairs) break;
            for (int64_t k = s; k < std::min(npairs, s + 10); k++) {
                IxOut r;
                const Poly &A = a->polys[pa[k]], &B = b->polys[pb[k]];
                bool bad = A.failed || B.failed;
                if (!bad) {
                    try {
                        r = intersection(A, B);
                    } catch (const RefAbort &) {
                        bad = true;
                    }
                }
                if (bad) {
                    r = IxOut();
                    r.n_inters = -3;
                    r.area = 0.0;
                    err = ORC_EREF_ABORT;
                }
                if (counts) {
                    counts[5 * k + 0] = r.n_inters;
                    counts[5 * k + 1] = r.n_segments;
                    counts[5 * k + 2] = r.n_in;
                    counts[5 * k + 3] = r.n_out;
                    counts[5 * k + 4] = r.n_border;
                }
                if (area) area[k] = r.area;
            }
        }
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    return err.load();
}

int orc_within(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa, const uint32_t *pb,
               double d_km, uint8_t *out, double *min_dist, int nthreads, int64_t *stats) {
    for (int64_t k = 0; k < npairs; k++)
        if (pa[k] >= a->polys.size() || pb[k] >= b->polys.size()) return ORC_EINVAL;
    std::atomic<int64_t> next(0);
    std::atomic<int> err(ORC_OK);
    std::mutex mu;
    QStats total;
    auto work = [&]() {
        QStats qs;
        for (;;) {
            int64_t s = next.fetch_add(10);
            if (s >= npairs) break;
            for (int64_t k = s; k < std::min(npairs, s + 10); k++) {
                double md = 100000.0;
                const Poly &S = a->polys[pa[k]], &T = b->polys[pb[k]];
                if (S.failed || T.failed) {
                    out[k] = 2;
                    err = ORC_EREF_ABORT;
                } else {
                    try {
                        out[k] = within(S, T, d_km, qs, &md) ? 1 : 0;
                    } catch (const RefAbort &) {
                        out[k] = 2;
                        err = ORC_EREF_ABORT;
                    }
                }
                if (min_dist) min_dist[k] = md;
            }
        }
        std::lock_guard<std::mutex> g(mu);
        total.bb_pairs += qs.bb_pairs;
        total.seg_tests += qs.seg_tests;
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    if (stats) {
        stats[0] = total.bb_pairs;
        stats[1] = total.seg_tests;
    }
    return err.load();
}

int64_t orc_candidates_within(const orc_set *s, double d_km, int64_t cap, uint32_t *pa, uint32_t *pb) {
    /* within_polygon_cpu rtree_query + MySearchCallback
     * (src/query/within_polygon_cpu.cpp:8-48): every (i, j), i != j, whose
     * MBBs overlap after expanding j's MBB by d_km; MBBs as loaded (before
     * the hierarchy's grid_align moves them) */
    int64_t n = 0;
    for (size_t j = 0; j < s->polys.size(); j++) {
        const Box q = box_expand(s->polys[j].mbb0, d_km);
        for (size_t i = 0; i < s->polys.size(); i++) {
            if (i == j || !box_intersect(s->polys[i].mbb0, q)) continue;
            if (n < cap) {
                pa[n] = (uint32_t)i;
                pb[n] = (uint32_t)j;
            }
            n++;
        }
    }
    return n;
}

int64_t orc_candidates(const orc_set *a, const orc_set *b, int64_t cap, uint32_t *pa, uint32_t *pb) {
    int64_t n = 0;
    for (size_t j = 0; j < b->polys.size(); j++)
        for (size_t i = 0; i < a->polys.size(); i++)
            if (box_intersect(a->polys[i].mbr, b->polys[j].mbr)) {
                if (n < cap) {
                    pa[n] = (uint32_t)i;
                    pb[n] = (uint32_t)j;
                }
                n++;
            }
    return n;
}

int orc_segment_intersect(const float *v) {
    Pt a = {v[0], v[1]}, b = {v[2], v[3]}, c = {v[4], v[5]}, d = {v[6], v[7]};
    return segment_intersect(a, b, c, d) ? 1 : 0;
}

int orc_double_to_int(double v) { return double_to_int(v); }

double orc_point_to_segment_distance(const float *p, const float *p1, const float *p2, int geography) {
    Pt a = {p[0], p[1]}, b = {p1[0], p1[1]}, c = {p2[0], p2[1]};
    return point_to_segment_distance(a, b, c, geography != 0);
}

} // extern "C"

extern "C" int orc_failed(const orc_set *s, int64_t i) {
    if (i < 0 || i >= (int64_t)s->polys.size()) return -1;
    return s->polys[i].failed ? 1 : 0;
}
