// Host raster builder of the IDEAL model (one polygon at a time).
#pragma once
#include <cstdint>
#include <vector>

namespace farm {

struct RasterGrid {
    double low[2], high[2];
    double step[2];
    int dimx = 0, dimy = 0;
};

struct RasterOut {
    RasterGrid g;
    std::vector<uint8_t> status;   // dimx*dimy fill classes (0 OUT, 15 IN)
    std::vector<double> areas;     // clipped polygon area per pixel
    std::vector<uint32_t> pixoff;  // dimx*dimy+1, relative seq index
    std::vector<uint32_t> seqs;    // (first edge, #edges) pairs
    std::vector<uint32_t> voff;    // dimx+2
    std::vector<double> vnodes;
    std::vector<uint32_t> hoff;    // dimy+2
    std::vector<double> hnodes;
    bool ok = true;                // false: the reference raster builder asserts here
};

// Scratch buffers reused across polygons by one worker thread.
struct RasterScratch;
RasterScratch *raster_scratch_new();
void raster_scratch_free(RasterScratch *s);

// ring: nv points (x,y interleaved) incl. the repeated first vertex;
// mbb: the loader's MBB (low x, low y, high x, high y).
void build_raster(const float *ring, int nv, const double mbb[4], int vpr, RasterOut &out, RasterScratch *s);

// Grid sizing of MyRaster::init_raster(int) for num_pixels pixels.
RasterGrid size_grid(const double mbb[4], int num_pixels);

// One coarse layer of the hierarchy (-h): its grid and fill classes.
struct LayerOut {
    RasterGrid g;
    std::vector<uint8_t> status;
};

// -h rasterisation: grid snapped to the universal 1.1e-5 * 2^k steps (square
// pixels, MBB floored/ceiled to the grid), the raster built on it, and the
// coarse layers (step doubled until dimx*dimy <= nlow) with merged fill
// classes.  layers is coarse to fine; its last entry is out's own grid (with
// an empty status: the finest layer's status is out.status).
void build_raster_hier(const float *ring, int nv, const double mbb[4], int vpr, int nlow, RasterOut &out,
                       std::vector<LayerOut> &layers, RasterScratch *s);

// The grids build_raster_hier uses: the snapped finest grid and the coarse
// layers (coarse to fine).  false when grid_align fails: the reference
// raster builder asserts on such a polygon.
bool hier_grids(const double mbb[4], int nv, int vpr, int nlow, RasterGrid &fine, std::vector<RasterGrid> &coarse);

}  // namespace farm

struct farm_hostset;
namespace farm {
// A host set over closed rings in the stored layout (each ring >= 3 stored
// vertices), taking the arrays over; MBBs computed on nthreads threads.
farm_hostset *hostset_adopt(std::vector<uint64_t> &&ring_off, std::vector<float> &&xy, int nthreads);
}  // namespace farm
