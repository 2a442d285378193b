// Device-side view of an uploaded polygon set and the shared predicates of
// the raster filter and segment refinement (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/farm_capi.h"

namespace farm {

constexpr int kCC = FARM_CATEGORY_COUNT;

struct DevSet {
    const farm_raster_hdr *hdr;
    const uint8_t *status;
    const uint32_t *pixoff;
    const uint2 *seqs;
    const float *x;
    const float *y;
    // per vertex v: (x[v], y[v], x[v] - x[v - 1], y[v] - y[v - 1]), the
    // differences rounded to f32, (0, 0) at v = 0 (refinement only: one
    // 16-byte load per target step)
    const float4 *xyd;
    // per edge sequence: index of its first C-step chunk box; the boxes
    // (centre x, y, half extents x, y), refinement only
    const uint32_t *choff;
    const float4 *chbox;
    // the same for chunks of kIxChunk segments, for the intersection's chunk
    // skip (the refinement's arrays when C == kIxChunk; null: no skip)
    const uint32_t *ixoff;
    const float4 *ixbox;
    const uint32_t *pixedges;  // per pixel: sum of sequence lengths (BORDER), else 0
    const uint32_t *voff;
    const double *vnodes;
    // fill classes for the approximate mode and the features (the status
    // bytes unless -w chose another category count) and that count
    const uint8_t *fill;
    int32_t cc;
    // hierarchy (-h) and per-pixel areas (within only; null otherwise)
    const farm_layer *layers;
    const uint8_t *lstatus;
    const double *areas;
};

// The refinement walks a target edge sequence in chunks of C steps (8, 16 or
// 32; kRefChunkDefault until the autotuner or the caller picks another) and
// screens kRefScreen chunk boxes per scalar load (the box array keeps
// kRefScreen spare entries past its end).
constexpr uint32_t kRefChunkDefault = 16;
// target segments per chunk of the intersection's collection screen
constexpr uint32_t kIxChunk = 16;
constexpr uint32_t kRefScreen = 4;

// A BORDER x BORDER pixel pair of one candidate pair.
struct PixPair {
    uint32_t pair;  // index into the current batch
    uint32_t pa;    // pixel id in the source raster
    uint32_t pb;    // pixel id in the target raster
    uint32_t slots; // edge slots of pa: sum over its sequences of (len + 1)
};

// util.h:98-106 double_to_int
__device__ __forceinline__ int d2i(double v) {
    int i = (int)v;
    if (fabs(1.0 * (i + 1) - v) < 0.00000001) i++;
    return i;
}

// MyRaster::get_offset_x/_y (MyRaster.cpp:188-200)
__device__ __forceinline__ int off_axis(double v, double lo, double step, int dim) {
    int i = d2i((v - lo) / step);
    return min(max(i, 0), dim - 1);
}

struct Window {
    int x0, x1, y0, y1;
};

// MyRaster::retrieve_pixels (MyRaster.cpp:332-346): pixel range of a box,
// shrunk by 1e-6 on every side.  shrink = 0 gives the unshrunk range the
// reference GPU filter takes over the target MBB (query_intersect.cu:51-54).
__device__ __forceinline__ Window window_of(const farm_raster_hdr &r, double lx, double ly, double hx, double hy,
                                            double shrink = 1e-6) {
    Window w;
    w.x0 = off_axis(lx + shrink, r.mbr[0], r.step[0], r.dimx);
    w.y0 = off_axis(ly + shrink, r.mbr[1], r.step[1], r.dimy);
    w.x1 = off_axis(hx - shrink, r.mbr[0], r.step[0], r.dimx);
    w.y1 = off_axis(hy - shrink, r.mbr[1], r.step[1], r.dimy);
    return w;
}

// MyRaster::get_pixel_box (MyRaster.cpp:312-322)
__device__ __forceinline__ void pixel_box(const farm_raster_hdr &r, int x, int y, double &lx, double &ly, double &hx,
                                          double &hy) {
    lx = r.mbr[0] + x * r.step[0];
    ly = r.mbr[1] + y * r.step[1];
    hx = r.mbr[0] + (x + 1) * r.step[0];
    hy = r.mbr[1] + (y + 1) * r.step[1];
}

// show_status (MyRaster.cpp:228-239): 0 OUT, 1 BORDER, 2 IN
__device__ __forceinline__ int show(uint8_t s) { return s == 0 ? 0 : (s == kCC - 1 ? 2 : 1); }

// Ideal::decodePixelArea (Ideal.cpp:47-62), midpoint of the fill interval
__device__ __forceinline__ double apx_area(uint8_t f, double pix_area) {
    double lo, hi;
    if (f == 0) {
        lo = hi = 0.0;
    } else if (f == kCC - 1) {
        lo = hi = pix_area;
    } else {
        lo = (1.0 * f - 1) / (kCC - 2) * pix_area;
        hi = (1.0 * f - 0) / (kCC - 2) * pix_area;
    }
    return (lo + hi) / 2;
}

// decodePixelArea at cc categories (gpu_decode_fullness, geometry.cuh:374-388,
// with the filter's category_count), midpoint of the fill interval
__device__ __forceinline__ double apx_area_cc(uint8_t f, double pix_area, int cc) {
    double lo, hi;
    if (f == 0) {
        lo = hi = 0.0;
    } else if (f == cc - 1) {
        lo = hi = pix_area;
    } else {
        lo = (1.0 * f - 1) / (cc - 2) * pix_area;
        hi = (1.0 * f - 0) / (cc - 2) * pix_area;
    }
    return (lo + hi) / 2;
}

// Point::cross(a, b) of the reference (Point.h:93-97), IEEE binary32 with
// every product and difference rounded separately (the build disables FMA
// contraction for device code).
__device__ __forceinline__ float cross3(float px, float py, float ax, float ay, float bx, float by) {
    float ux = ax - px, uy = ay - py;
    float vx = bx - px, vy = by - py;
    float l = ux * vy;
    float r = uy * vx;
    return l - r;
}

__device__ __forceinline__ int sgnf(float x) { return (x > 0.f) - (x < 0.f); }

// inter1 (geometry_computation.h:195-212)
__device__ __forceinline__ bool inter1(float a, float b, float c, float d) {
    float t;
    if (a > b) { t = a; a = b; b = t; }
    if (c > d) { t = c; c = d; d = t; }
    return fmaxf(a, c) <= fminf(b, d);
}

// segment_intersect (geometry_computation.h:215-221) given the four
// orientations o1 = a.cross(b,c), o2 = a.cross(b,d), o3 = c.cross(a,d),
// o4 = c.cross(b,d).  c.cross(d,a) and c.cross(d,b) are exactly -o3 and -o4
// (same rounded products, operands swapped), so their signs compare like
// those of o3 and o4.
__device__ __forceinline__ bool seg_hit(float o1, float o2, float o3, float o4, float ax, float ay, float bx, float by,
                                        float cx, float cy, float dx, float dy) {
    if (o3 == 0.f && o4 == 0.f) return inter1(ax, bx, cx, dx) && inter1(ay, by, cy, dy);
    return sgnf(o1) != sgnf(o2) && sgnf(o3) != sgnf(o4);
}

__device__ __forceinline__ bool seg_intersect(float ax, float ay, float bx, float by, float cx, float cy, float dx,
                                              float dy) {
    float o3 = cross3(cx, cy, ax, ay, dx, dy);
    float o4 = cross3(cx, cy, bx, by, dx, dy);
    float o1 = cross3(ax, ay, bx, by, cx, cy);
    float o2 = cross3(ax, ay, bx, by, dx, dy);
    return seg_hit(o1, o2, o3, o4, ax, ay, bx, by, cx, cy, dx, dy);
}

}  // namespace farm
