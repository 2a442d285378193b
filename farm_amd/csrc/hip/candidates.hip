// Candidate pairs on gfx950: the MBB join that feeds the queries.
//
// Replaces the host R-tree probe rtree_query / MySearchCallback
// (src/query/intersect_cpu.cpp:7-36, src/query/within_polygon_cpu.cpp:8-48,
// RTree::Search src/index/RTree.h:581-600): every (i, j) whose boxes overlap
// inclusively (box::intersect), the target box expanded by d km first for a
// within join (box::expand, Box.cpp:200-214), i == j skipped in a self-join.
// The result is the same set, ordered by target then source, which is the
// order farm_candidates returns; it stays on the device as the handle's pair
// list, ready for the queries.
//
// A uniform grid over the source boxes:
//  1. every source box is listed in each cell it covers (count, scan, fill)
//  2. the probe: one lane per unit, a unit being one target box and one cell
//     it covers (a large target spreads over many lanes instead of holding
//     one lane for all its cells); the lane tests the cell's listed sources,
//     and a pair is reported only from the cell where the two boxes' cell
//     ranges first meet (the larger low cell on each axis), so each
//     overlapping pair is counted once (pass 1) and written once (pass 2) as
//     a key target << 32 | source
//  3. a radix sort of the keys gives the order; the keys are split into the
//     pair arrays.
// The boxes are compared in double exactly as the host does; cell indexes
// only choose which sources a lane visits, and both the listing and the
// visiting use the same device function, so every overlapping pair's cell
// ranges meet (the cell index is monotone in the coordinate).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "handle.h"

namespace farm {

void expand_box_km(double q[4], double km);  // host: box::expand, shared with farm_candidates

namespace {

struct Grid {
    double x0, y0, ix, iy;  // cell of v: floor((v - x0) * ix), clamped
    int nx, ny;
};

__host__ __device__ __forceinline__ int cell_of(double v, double lo, double inv, int n) {
    double t = floor((v - lo) * inv);
    t = fmin(fmax(t, 0.0), (double)(n - 1));
    return (int)t;
}

// box::intersect (inclusive), a = source, q = target
__device__ __forceinline__ bool boxes_overlap(const double *a, const double *q) {
    return !(a[0] > q[2] || q[0] > a[2] || a[1] > q[3] || q[1] > a[3]);
}

struct CellRange {
    int x0, y0, x1, y1;
};

__device__ __forceinline__ CellRange cells_of(const Grid &g, const double *b) {
    return {cell_of(b[0], g.x0, g.ix, g.nx), cell_of(b[1], g.y0, g.iy, g.ny), cell_of(b[2], g.x0, g.ix, g.nx),
            cell_of(b[3], g.y0, g.iy, g.ny)};
}

__global__ __launch_bounds__(256) void k_cg_list_count(const double *__restrict__ boxa, uint32_t na, Grid g,
                                                       uint32_t *__restrict__ cellcnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const CellRange r = cells_of(g, boxa + 4ull * i);
    for (int y = r.y0; y <= r.y1; y++)
        for (int x = r.x0; x <= r.x1; x++) atomicAdd(&cellcnt[(uint32_t)y * g.nx + x], 1u);
}

__global__ __launch_bounds__(256) void k_cg_list_fill(const double *__restrict__ boxa, uint32_t na, Grid g,
                                                      uint32_t *__restrict__ cursor, uint32_t *__restrict__ items) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const CellRange r = cells_of(g, boxa + 4ull * i);
    for (int y = r.y0; y <= r.y1; y++)
        for (int x = r.x0; x <= r.x1; x++) items[atomicAdd(&cursor[(uint32_t)y * g.nx + x], 1u)] = i;
}

// Probe units of each target box: the cells its box covers.
__global__ __launch_bounds__(256) void k_cg_units(const double *__restrict__ boxq, uint32_t nb, Grid g,
                                                  uint64_t *__restrict__ ucnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const CellRange r = cells_of(g, boxq + 4ull * j);
    ucnt[j] = (uint64_t)(r.x1 - r.x0 + 1) * (uint64_t)(r.y1 - r.y0 + 1);
}

// One lane per probe unit u (targets in order, each target's cells row by
// row; uoff: the targets' first units).  WRITE = false: the unit's pair count
// into cnt[u]; true: its keys from off[u] on.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_cg_probe(const double *__restrict__ boxa, const double *__restrict__ boxq,
                                                  uint32_t nb, const uint64_t *__restrict__ uoff, uint64_t nunits,
                                                  Grid g, const uint32_t *__restrict__ celloff,
                                                  const uint32_t *__restrict__ items, int skip_self,
                                                  uint64_t *__restrict__ cnt, const uint64_t *__restrict__ off,
                                                  unsigned long long *__restrict__ keys) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    uint32_t lo = 0, hi = nb;  // the target owning u: the last j with uoff[j] <= u
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uoff[mid] <= u) lo = mid;
        else hi = mid;
    }
    const uint32_t j = lo;
    const double q[4] = {boxq[4ull * j], boxq[4ull * j + 1], boxq[4ull * j + 2], boxq[4ull * j + 3]};
    const CellRange r = cells_of(g, q);
    const uint64_t k = u - uoff[j];
    const uint64_t wx = (uint64_t)(r.x1 - r.x0 + 1);
    const int x = r.x0 + (int)(k % wx), y = r.y0 + (int)(k / wx);
    const uint32_t c = (uint32_t)y * g.nx + x;
    uint64_t n = 0;
    uint64_t pos = WRITE ? off[u] : 0;
    for (uint32_t t = celloff[c]; t < celloff[c + 1]; t++) {
        const uint32_t i = items[t];
        if (skip_self && i == j) continue;
        const double *a = boxa + 4ull * i;
        if (!boxes_overlap(a, q)) continue;
        const CellRange ra = cells_of(g, a);
        if (max(ra.x0, r.x0) != x || max(ra.y0, r.y0) != y) continue;  // reported from another cell
        if (WRITE) keys[pos++] = ((unsigned long long)j << 32) | i;
        else n++;
    }
    if (!WRITE) cnt[u] = n;
}

__global__ __launch_bounds__(256) void k_cg_split(const unsigned long long *__restrict__ keys, uint64_t n,
                                                  uint32_t a_base, uint32_t b_base, uint32_t *__restrict__ pa,
                                                  uint32_t *__restrict__ pb) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long v = keys[k];
    pa[k] = a_base + (uint32_t)(v & 0xffffffffull);
    pb[k] = b_base + (uint32_t)(v >> 32);
}

// Grid over the source boxes: cells about the median source box in size (at
// most 4096 per axis), coarsened while the listing would exceed 16 entries
// per source (a few huge boxes over a fine grid) or the probe 2^28 units.
Grid make_grid(const std::vector<double> &a, int64_t na, const std::vector<double> &q, int64_t nb) {
    Grid g;
    double ux0 = INFINITY, uy0 = INFINITY, ux1 = -INFINITY, uy1 = -INFINITY;
    std::vector<double> w(na), hgt(na);
    for (int64_t i = 0; i < na; i++) {
        const double *b = &a[4 * i];
        ux0 = std::min(ux0, b[0]);
        uy0 = std::min(uy0, b[1]);
        ux1 = std::max(ux1, b[2]);
        uy1 = std::max(uy1, b[3]);
        w[i] = b[2] - b[0];
        hgt[i] = b[3] - b[1];
    }
    std::nth_element(w.begin(), w.begin() + na / 2, w.end());
    std::nth_element(hgt.begin(), hgt.begin() + na / 2, hgt.end());
    const double ex = ux1 - ux0, ey = uy1 - uy0;
    auto dim = [](double extent, double cell) {
        if (!(extent > 0) || !(cell > 0)) return 1;
        return (int)std::min(4096.0, std::max(1.0, std::ceil(extent / cell)));
    };
    int nx = dim(ex, std::max(w[na / 2], ex / 4096)), ny = dim(ey, std::max(hgt[na / 2], ey / 4096));
    for (;;) {
        g.x0 = ux0;
        g.y0 = uy0;
        g.nx = nx;
        g.ny = ny;
        g.ix = ex > 0 ? nx / ex : 0.0;
        g.iy = ey > 0 ? ny / ey : 0.0;
        // the listing size estimated with the same cell function
        uint64_t entries = 0;
        for (int64_t i = 0; i < na; i++) {
            const double *b = &a[4 * i];
            entries += (uint64_t)(cell_of(b[2], g.x0, g.ix, nx) - cell_of(b[0], g.x0, g.ix, nx) + 1) *
                       (uint64_t)(cell_of(b[3], g.y0, g.iy, ny) - cell_of(b[1], g.y0, g.iy, ny) + 1);
        }
        uint64_t units = 0;
        for (int64_t j = 0; j < nb; j++) {
            const double *b = &q[4 * j];
            units += (uint64_t)(cell_of(b[2], g.x0, g.ix, nx) - cell_of(b[0], g.x0, g.ix, nx) + 1) *
                     (uint64_t)(cell_of(b[3], g.y0, g.iy, ny) - cell_of(b[1], g.y0, g.iy, ny) + 1);
        }
        if ((entries <= 16ull * (uint64_t)na + (1ull << 20) && units <= (1ull << 28)) || (nx == 1 && ny == 1))
            break;
        nx = std::max(1, nx / 2);
        ny = std::max(1, ny / 2);
    }
    return g;
}

}  // namespace
}  // namespace farm

using namespace farm;

extern "C" {

int farm_generate_pairs(farm_handle h, const double *mbb_a, int64_t na, const double *mbb_b, int64_t nb,
                        double expand_km, int skip_self, uint32_t a_base, uint32_t b_base, uint64_t *npairs) {
    if (!h || na < 0 || nb < 0 || (na && !mbb_a) || (nb && !mbb_b) || !(expand_km >= 0))
        return fail(h, FARM_EINVAL, "bad generate_pairs arguments");
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "upload polygons before pairs");
    if ((int64_t)a_base + na > h->npolys || (int64_t)b_base + nb > h->npolys)
        return fail(h, FARM_EINVAL, "box sets reach past the %lld uploaded polygons", (long long)h->npolys);
    if (na >= (1ll << 31) - 1 || nb >= (1ll << 31) - 1) return fail(h, FARM_ECAPACITY, "2^31 boxes or more");
    auto bad_box = [](const double *b) {
        for (int k = 0; k < 4; k++)
            if (!std::isfinite(b[k])) return true;
        return b[0] > b[2] || b[1] > b[3];
    };
    for (int64_t i = 0; i < na; i++)
        if (bad_box(mbb_a + 4 * i)) return fail(h, FARM_EINVAL, "source box %lld is not finite or inverted", (long long)i);
    for (int64_t j = 0; j < nb; j++)
        if (bad_box(mbb_b + 4 * j)) return fail(h, FARM_EINVAL, "target box %lld is not finite or inverted", (long long)j);
    HIPCHK(h, hipSetDevice(h->device));
    if (na == 0 || nb == 0) {
        h->npairs = 0;
        if (npairs) *npairs = 0;
        return FARM_OK;
    }
    std::vector<double> a(mbb_a, mbb_a + 4 * na), q(mbb_b, mbb_b + 4 * nb);
    if (expand_km > 0)
        for (int64_t j = 0; j < nb; j++) expand_box_km(&q[4 * j], expand_km);
    const Grid g = make_grid(a, na, q, nb);
    const uint64_t ncell = (uint64_t)g.nx * g.ny;
    int rc;
    if ((rc = upload(h, h->cg_boxa, a.data(), 32ull * na))) return rc;
    if ((rc = upload(h, h->cg_boxq, q.data(), 32ull * nb))) return rc;
    // 1. the listing
    if ((rc = ensure(h, h->cg_cellcnt, 4 * (ncell + 1)))) return rc;
    if ((rc = ensure(h, h->cg_celloff, 4 * (ncell + 1)))) return rc;
    if ((rc = fill_u32(h, h->cg_cellcnt.p, ncell + 1, 0))) return rc;
    hipLaunchKernelGGL(k_cg_list_count, dim3(blocks_for(na)), dim3(256), 0, h->stream, P<double>(h->cg_boxa),
                       (uint32_t)na, g, P<uint32_t>(h->cg_cellcnt));
    HIPCHK(h, hipGetLastError());
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, P<uint32_t>(h->cg_cellcnt), P<uint32_t>(h->cg_celloff),
                                               (int)(ncell + 1), h->stream));
    if ((rc = ensure(h, h->scan_tmp, tmp))) return rc;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp.p, tmp, P<uint32_t>(h->cg_cellcnt),
                                               P<uint32_t>(h->cg_celloff), (int)(ncell + 1), h->stream));
    uint32_t nitems = 0;
    HIPCHK(h, hipMemcpyAsync(&nitems, P<uint32_t>(h->cg_celloff) + ncell, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if ((rc = ensure(h, h->cg_items, 4ull * nitems))) return rc;
    // the cursor reuses the count buffer
    HIPCHK(h, hipMemcpyAsync(h->cg_cellcnt.p, h->cg_celloff.p, 4 * ncell, hipMemcpyDeviceToDevice, h->stream));
    hipLaunchKernelGGL(k_cg_list_fill, dim3(blocks_for(na)), dim3(256), 0, h->stream, P<double>(h->cg_boxa),
                       (uint32_t)na, g, P<uint32_t>(h->cg_cellcnt), P<uint32_t>(h->cg_items));
    HIPCHK(h, hipGetLastError());
    // 2. probe units, count, scan, write
    if ((rc = ensure(h, h->cg_uoff, 16ull * (nb + 1)))) return rc;  // unit counts, then their offsets
    uint64_t *ucnt = P<uint64_t>(h->cg_uoff), *uoff = ucnt + (nb + 1);
    if ((rc = fill_u32(h, ucnt, 2 * (nb + 1), 0))) return rc;
    hipLaunchKernelGGL(k_cg_units, dim3(blocks_for(nb)), dim3(256), 0, h->stream, P<double>(h->cg_boxq),
                       (uint32_t)nb, g, ucnt);
    HIPCHK(h, hipGetLastError());
    tmp = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ucnt, uoff, (int)(nb + 1), h->stream));
    if ((rc = ensure(h, h->scan_tmp, tmp))) return rc;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp.p, tmp, ucnt, uoff, (int)(nb + 1), h->stream));
    uint64_t nunits = 0;
    HIPCHK(h, hipMemcpyAsync(&nunits, uoff + nb, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (nunits >= (1ull << 31)) return fail(h, FARM_ECAPACITY, "probe of %llu units", (unsigned long long)nunits);
    if ((rc = ensure(h, h->cg_cnt, 8ull * (nunits + 1)))) return rc;
    if ((rc = ensure(h, h->cg_off, 8ull * (nunits + 1)))) return rc;
    if ((rc = fill_u32(h, h->cg_cnt.p, 2 * (nunits + 1), 0))) return rc;
    const unsigned pblocks = blocks_for(nunits);
    hipLaunchKernelGGL(k_cg_probe<false>, dim3(pblocks), dim3(256), 0, h->stream, P<double>(h->cg_boxa),
                       P<double>(h->cg_boxq), (uint32_t)nb, (const uint64_t *)uoff, nunits, g,
                       P<uint32_t>(h->cg_celloff), P<uint32_t>(h->cg_items), skip_self, P<uint64_t>(h->cg_cnt),
                       (const uint64_t *)nullptr, (unsigned long long *)nullptr);
    HIPCHK(h, hipGetLastError());
    tmp = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, P<uint64_t>(h->cg_cnt), P<uint64_t>(h->cg_off),
                                               (int)(nunits + 1), h->stream));
    if ((rc = ensure(h, h->scan_tmp, tmp))) return rc;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp.p, tmp, P<uint64_t>(h->cg_cnt), P<uint64_t>(h->cg_off),
                                               (int)(nunits + 1), h->stream));
    uint64_t total = 0;
    HIPCHK(h, hipMemcpyAsync(&total, P<uint64_t>(h->cg_off) + nunits, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (total >= (1ull << 31)) return fail(h, FARM_ECAPACITY, "%llu pairs: 2^31 or more", (unsigned long long)total);
    if ((rc = ensure(h, h->cg_keys, 8 * total))) return rc;
    if ((rc = ensure(h, h->cg_keys2, 8 * total))) return rc;
    hipLaunchKernelGGL(k_cg_probe<true>, dim3(pblocks), dim3(256), 0, h->stream, P<double>(h->cg_boxa),
                       P<double>(h->cg_boxq), (uint32_t)nb, (const uint64_t *)uoff, nunits, g,
                       P<uint32_t>(h->cg_celloff), P<uint32_t>(h->cg_items), skip_self, (uint64_t *)nullptr,
                       P<uint64_t>(h->cg_off), P<unsigned long long>(h->cg_keys));
    HIPCHK(h, hipGetLastError());
    // 3. order by (target, source) and split
    int end_bit = 32;
    while (end_bit < 64 && ((uint64_t)(nb - 1) >> (end_bit - 32)) != 0) end_bit++;
    tmp = 0;
    HIPCHK(h, hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, P<unsigned long long>(h->cg_keys),
                                                P<unsigned long long>(h->cg_keys2), (int)total, 0, end_bit,
                                                h->stream));
    if ((rc = ensure(h, h->scan_tmp, tmp))) return rc;
    HIPCHK(h, hipcub::DeviceRadixSort::SortKeys(h->scan_tmp.p, tmp, P<unsigned long long>(h->cg_keys),
                                                P<unsigned long long>(h->cg_keys2), (int)total, 0, end_bit,
                                                h->stream));
    if ((rc = ensure(h, h->pa, 4 * total))) return rc;
    if ((rc = ensure(h, h->pb, 4 * total))) return rc;
    if (total)
        hipLaunchKernelGGL(k_cg_split, dim3(blocks_for(total)), dim3(256), 0, h->stream,
                           P<unsigned long long>(h->cg_keys2), total, a_base, b_base, P<uint32_t>(h->pa),
                           P<uint32_t>(h->pb));
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->npairs = total;
    if (npairs) *npairs = total;
    return FARM_OK;
}

int farm_download_pairs(farm_handle h, uint64_t begin, uint64_t end, uint32_t *a, uint32_t *b) {
    if (!h || begin > end || (end > begin && (!a || !b))) return fail(h, FARM_EINVAL, "bad download arguments");
    if (end > h->npairs)
        return fail(h, FARM_EINVAL, "pair window [%llu,%llu) outside [0,%llu)", (unsigned long long)begin,
                    (unsigned long long)end, (unsigned long long)h->npairs);
    if (end == begin) return FARM_OK;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipMemcpyAsync(a, P<uint32_t>(h->pa) + begin, 4 * (end - begin), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(b, P<uint32_t>(h->pb) + begin, 4 * (end - begin), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

}  // extern "C"
