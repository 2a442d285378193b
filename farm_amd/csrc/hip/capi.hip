// C ABI of libfarm.so: device memory, uploads and the intersect driver.
// Declarations and their reference counterparts: include/farm_capi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <string>
#include <vector>

#include "handle.h"
#include "isect_work.h"

namespace farm {
__global__ void k_filter(DevSet, const uint32_t *, const uint32_t *, const IsxDesc *, int, int, const uint8_t *,
                         uint32_t *, const uint32_t *, PixPair *, uint64_t, uint32_t *, unsigned long long *);
__global__ void k_scan_sums(const uint32_t *, const unsigned long long *, uint64_t, uint32_t *);
__global__ void k_scan_parts(uint32_t *, const unsigned long long *, uint64_t, uint32_t *);
__global__ void k_scan_apply(const uint32_t *, const unsigned long long *, uint64_t, const uint32_t *, uint32_t *);
__global__ void k_set_desc(IsxDesc *, unsigned long long, unsigned long long, uint8_t *, uint32_t *);
__global__ void k_bucket_hist(DevSet, const uint32_t *, const IsxDesc *, const PixPair *, const unsigned long long *,
                              uint64_t, const float *, uint8_t *, uint32_t *);
__global__ void k_bucket_scan(uint32_t *, uint32_t *);
__global__ void k_bucket_scatter(const uint8_t *, const unsigned long long *, uint64_t, uint32_t *, uint32_t *);
template <int E>
__global__ void k_tiles(const PixPair *, const uint32_t *, unsigned long long *, uint64_t, uint32_t *);
template <int E, int C>
__global__ void k_refine(DevSet, const uint32_t *, const uint32_t *, const IsxDesc *, const TileDesc *, uint64_t,
                         uint64_t, uint32_t *, uint32_t *, unsigned long long *, int);
__global__ void k_tile_map(DevSet, const uint32_t *, const uint32_t *, const IsxDesc *, const PixPair *,
                           const uint32_t *, const uint32_t *, const unsigned long long *, uint64_t, uint64_t,
                           TileDesc *);
__global__ void k_plain(DevSet, const uint32_t *, const uint32_t *, uint32_t, uint32_t *, unsigned long long *);
__global__ void k_pixedges(DevSet, int64_t, uint32_t *);
__global__ void k_vdiff(const float *, const float *, uint64_t, float4 *);
__global__ void k_chunk_count(DevSet, int64_t, uint32_t, uint32_t *);
__global__ void k_chunk_box(DevSet, int64_t, uint32_t, const uint32_t *, float4 *);
__global__ void k_finalize(const uint32_t *, const IsxDesc *, unsigned long long *);
}  // namespace farm

extern "C" int farm_mlp_predict_pairs(farm_handle h, uint64_t begin, uint64_t end, float *out);

using namespace farm;

namespace farm {
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Plain streaming copy (16 B per lane per step): the reference point for
// HBM byte counters and bandwidth.
__global__ __launch_bounds__(256) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                    uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// learned -a decision: an IN overlap decides true, no BORDER pair decides
// false, otherwise the intersect logit: > accept -> true, < reject -> false,
// in between -> exact refinement (skip = 0)
__global__ __launch_bounds__(256) void k_learned_decide(const int8_t *__restrict__ decided,
                                                         const float *__restrict__ pred, uint32_t n, float accept,
                                                         float reject, uint32_t *__restrict__ result,
                                                         uint8_t *__restrict__ skip) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = decided[i];
    const float z = pred[2 * (size_t)i];
    uint32_t r = 0;
    uint8_t s = 1;
    if (d == 1) r = 1;
    else if (d == 0) r = 0;
    else if (z > accept) r = 1;
    else if (z < reject) r = 0;
    else s = 0;
    result[i] = r;
    skip[i] = s;
}
}  // namespace farm

namespace {

template <class T>
int exclusive_scan(farm_handle h, const T *in, T *out, uint64_t n) {
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, h->stream));
    int rc = ensure(h, h->scan_tmp, tmp);
    if (rc) return rc;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp.p, tmp, in, out, (int)n, h->stream));
    return FARM_OK;
}

// Exclusive scan on the device of n = bounded_count(nsrc, cap) values,
// out[n] = total (no host round trip, no temporaries beyond W.part).
int scan_dev(farm_handle h, const IsxWork &W, const uint32_t *in, const unsigned long long *nsrc, uint64_t cap,
             uint32_t *out) {
    hipLaunchKernelGGL(k_scan_sums, dim3(kScanParts), dim3(256), 0, h->stream, in, nsrc, cap, W.part);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanParts), 0, h->stream, W.part, nsrc, cap, out);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(k_scan_apply, dim3(kScanParts), dim3(256), 0, h->stream, in, nsrc, cap,
                       (const uint32_t *)W.part, out);
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

// What one launch sequence of the batched intersect depends on (the key of
// its captured graph).  Everything else -- the pair window, the output and
// cost-label pointers -- is read from the workspace's IsxDesc.
struct IsxCfg {
    int32_t approx;           // reference -a: filter only
    int32_t tile;             // source edges per lane of the refinement
    const float *pred_cost;   // learned queue order: the cost head's predictions
    const uint8_t *skip;      // pairs decided before the filter (learned -a)
    const void *ws;           // workspace base and capacities
    uint64_t cap_pairs, cap_pp, cap_tiles;
    uint64_t set_gen;         // device arrays of the uploaded set and pairs
    const uint32_t *pa, *pb;  // the pair list
    int32_t grid;             // blocks of the grid-stride kernels
    int32_t prune;            // the refinement's exact chunk prune on / off
    int32_t chunk;            // target segments per chunk box of the refinement
    int32_t bpc;              // refinement workgroups per CU
};
static_assert(sizeof(IsxCfg) <= sizeof(((farm_handle_s::IsxGraph *)nullptr)->key), "graph key too large");

// The batched pipeline up to the refinement: filter count pass, scan, emit
// pass, cost order, tiles, tile map (exact), or the filter alone (-a).
int isx_front(farm_handle h, const DevSet &S, const IsxWork &W, const IsxCfg &c) {
    int rc;
    const unsigned g = (unsigned)c.grid;
    if ((rc = fill_u32(h, W.ctr, 2 * kCtrSlots, 0))) return rc;
    hipLaunchKernelGGL(k_filter, dim3(g), dim3(256), 0, h->stream, S, c.pa, c.pb, (const IsxDesc *)W.desc, 0,
                       c.approx, c.skip, W.npp, (const uint32_t *)nullptr, (PixPair *)nullptr, W.cap_pp, W.result,
                       W.ctr);
    HIPCHK(h, hipGetLastError());
    if (c.approx) return FARM_OK;
    if ((rc = scan_dev(h, W, W.npp, &W.desc->n, W.cap_pairs, W.pp_off))) return rc;
    hipLaunchKernelGGL(k_filter, dim3(g), dim3(256), 0, h->stream, S, c.pa, c.pb, (const IsxDesc *)W.desc, 1, 0,
                       c.skip, W.npp, (const uint32_t *)W.pp_off, W.pps, W.cap_pp, W.result, W.ctr);
    HIPCHK(h, hipGetLastError());
    if ((rc = fill_u32(h, W.hist, 64, 0))) return rc;
    hipLaunchKernelGGL(k_bucket_hist, dim3(g), dim3(256), 0, h->stream, S, c.pb, (const IsxDesc *)W.desc,
                       (const PixPair *)W.pps, (const unsigned long long *)W.ctr, W.cap_pp, c.pred_cost, W.bucket,
                       W.hist);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(64), 0, h->stream, W.hist, W.cursor);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(k_bucket_scatter, dim3(g), dim3(256), 0, h->stream, (const uint8_t *)W.bucket,
                       (const unsigned long long *)W.ctr, W.cap_pp, W.cursor, W.order);
    HIPCHK(h, hipGetLastError());
    switch (c.tile) {
        case 1:
            hipLaunchKernelGGL(k_tiles<1>, dim3(g), dim3(256), 0, h->stream, (const PixPair *)W.pps,
                               (const uint32_t *)W.order, W.ctr, W.cap_pp, W.tiles);
            break;
        case 2:
            hipLaunchKernelGGL(k_tiles<2>, dim3(g), dim3(256), 0, h->stream, (const PixPair *)W.pps,
                               (const uint32_t *)W.order, W.ctr, W.cap_pp, W.tiles);
            break;
        case 4:
            hipLaunchKernelGGL(k_tiles<4>, dim3(g), dim3(256), 0, h->stream, (const PixPair *)W.pps,
                               (const uint32_t *)W.order, W.ctr, W.cap_pp, W.tiles);
            break;
        default:
            hipLaunchKernelGGL(k_tiles<8>, dim3(g), dim3(256), 0, h->stream, (const PixPair *)W.pps,
                               (const uint32_t *)W.order, W.ctr, W.cap_pp, W.tiles);
            break;
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = scan_dev(h, W, W.tiles, &W.ctr[kCtrPixPairs], W.cap_pp, W.tile_off))) return rc;
    hipLaunchKernelGGL(k_tile_map, dim3(g), dim3(256), 0, h->stream, S, c.pa, c.pb, (const IsxDesc *)W.desc,
                       (const PixPair *)W.pps, (const uint32_t *)W.order, (const uint32_t *)W.tile_off,
                       (const unsigned long long *)W.ctr, W.cap_pp, W.cap_tiles, W.tile_map);
    HIPCHK(h, hipGetLastError());
    return fill_u32(h, W.cursor + 63, 1, 0);  // the refinement's tile cursor
}

// The persistent refinement with the tile width and chunk size of the call.
template <int E, int C>
void launch_refine(farm_handle h, const DevSet &S, const IsxWork &W, const IsxCfg &c) {
    const dim3 grid((unsigned)(h->num_cus * c.bpc)), blk(256);
    hipLaunchKernelGGL((k_refine<E, C>), grid, blk, 0, h->stream, S, c.pa, c.pb, (const IsxDesc *)W.desc,
                       (const TileDesc *)W.tile_map, W.cap_pp, W.cap_tiles, W.cursor + 63, W.result, W.ctr,
                       c.prune);
}

template <int E>
void launch_refine_chunk(farm_handle h, const DevSet &S, const IsxWork &W, const IsxCfg &c) {
    switch (c.chunk) {
        case 8:
            launch_refine<E, 8>(h, S, W, c);
            break;
        case 32:
            launch_refine<E, 32>(h, S, W, c);
            break;
        default:
            launch_refine<E, 16>(h, S, W, c);
            break;
    }
}

int isx_refine(farm_handle h, const DevSet &S, const IsxWork &W, const IsxCfg &c) {
    switch (c.tile) {
        case 1:
            launch_refine_chunk<1>(h, S, W, c);
            break;
        case 2:
            launch_refine_chunk<2>(h, S, W, c);
            break;
        case 4:
            launch_refine_chunk<4>(h, S, W, c);
            break;
        default:
            launch_refine_chunk<8>(h, S, W, c);
            break;
    }
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

void drop_graphs(farm_handle h) {
    for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

// The front of the pipeline as one graph launch: captured on the first
// call with this configuration, replayed afterwards.
int isx_front_graph(farm_handle h, const DevSet &S, const IsxWork &W, const IsxCfg &c) {
    unsigned char key[sizeof(((farm_handle_s::IsxGraph *)nullptr)->key)];
    memset(key, 0, sizeof key);
    memcpy(key, &c, sizeof c);
    for (auto &g : h->graphs)
        if (!memcmp(g.key, key, sizeof key)) {
            HIPCHK(h, hipGraphLaunch(g.exec, h->stream));
            return FARM_OK;
        }
    if (h->graphs.size() >= 16) drop_graphs(h);
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = isx_front(h, S, W, c);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(h, FARM_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    farm_handle_s::IsxGraph g;
    memcpy(g.key, key, sizeof key);
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(h, FARM_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
    h->graphs.push_back(g);
    HIPCHK(h, hipGraphLaunch(g.exec, h->stream));
    return FARM_OK;
}

// A workspace holding at least these capacities (the handle's own, grown:
// every captured graph refers to the old block and is dropped with it).
int isx_reserve(farm_handle h, uint64_t pairs, IsxWork *W) {
    const uint64_t cp = std::max<uint64_t>(h->isx_cap_pairs, std::max<uint64_t>(pairs, 1024));
    const uint64_t cq = std::min<uint64_t>(0xffffffffull,
                                           std::max<uint64_t>(h->isx_cap_pp, std::max(h->isx_want_pp, 4 * cp)));
    const uint64_t ct = std::min<uint64_t>(0xffffffffull,
                                           std::max<uint64_t>(h->isx_cap_tiles, std::max(h->isx_want_tiles, cq)));
    if (!h->isx.p || cp > h->isx_cap_pairs || cq > h->isx_cap_pp || ct > h->isx_cap_tiles) {
        drop_graphs(h);
        HIPCHK(h, hipStreamSynchronize(h->stream));
        free_buf(h->isx);
        h->isx_cap_pairs = h->isx_cap_pp = h->isx_cap_tiles = 0;
        int rc = ensure(h, h->isx, isx_bytes(cp, cq, ct));
        if (rc) return rc;
        h->isx_cap_pairs = cp;
        h->isx_cap_pp = cq;
        h->isx_cap_tiles = ct;
    }
    *W = isx_carve(h->isx.p, h->isx_cap_pairs, h->isx_cap_pp, h->isx_cap_tiles);
    return FARM_OK;
}

// Fixed scratch of farm_intersect_noalloc: at most kNaPairs pairs,
// kNaPixPairs BORDER x BORDER pixel pairs and kNaTiles refinement tiles per
// call (about 80 MiB, held by every GPU the library runs on until the
// process ends).  Host code never passes the array's name to a kernel: it
// uses the address hipGetSymbolAddress returns for the current device.
constexpr uint64_t kNaPairs = 1ull << 18, kNaPixPairs = 1ull << 20, kNaTiles = 1ull << 20;
constexpr uint64_t kNaBytes = isx_bytes(kNaPairs, kNaPixPairs, kNaTiles);

}  // namespace

namespace farm {
__device__ __align__(256) unsigned char scratch[kNaBytes];
}  // namespace farm

namespace farm {

int build_chunk_boxes(farm_handle h, uint32_t C);

// Sum and count of the nonzero per-pixel edge counts (BORDER pixels).
__global__ __launch_bounds__(256) void k_shape_stats(const uint32_t *__restrict__ pixedges, uint64_t n,
                                                     uint64_t *__restrict__ out) {
    unsigned long long sum = 0, cnt = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = pixedges[i];
        sum += e;
        cnt += e ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        cnt += __shfl_xor(cnt, o);
    }
    if ((threadIdx.x & 63) == 0 && cnt) {
        atomicAdd((unsigned long long *)&out[0], sum);
        atomicAdd((unsigned long long *)&out[1], cnt);
    }
}

// Everything the queries derive from an uploaded or device-built set: the
// vertex padding, per-vertex step differences, per-pixel edge counts, the
// set's shape bucket and the refinement's chunk boxes.  Synchronises the
// stream.
int finish_set(farm_handle h) {
    int rc;
    const uint64_t nverts = h->nverts;
    for (DevBuf *b : {&h->x, &h->y}) {
        if ((rc = ensure(h, *b, 4 * (nverts + kVertexPad)))) return rc;
        HIPCHK(h, hipMemsetAsync((char *)b->p + 4 * nverts, 0, 4 * kVertexPad, h->stream));
    }
    // successive vertex differences over the padded arrays (the last one
    // reads no vertex past the padding and is 0)
    if ((rc = ensure(h, h->xyd, 16 * (nverts + kVertexPad)))) return rc;
    hipLaunchKernelGGL(k_vdiff, dim3(blocks_for(nverts + kVertexPad)), dim3(256), 0, h->stream, P<float>(h->x),
                       P<float>(h->y), nverts + kVertexPad, P<float4>(h->xyd));
    HIPCHK(h, hipGetLastError());
    if ((rc = ensure(h, h->pixedges, 4 * h->nstatus))) return rc;
    h->ring_area_ready = false;
    h->set_gen++;
    if (h->npolys > 0) {
        hipLaunchKernelGGL(k_pixedges, dim3((unsigned)h->npolys), dim3(256), 0, h->stream, devset(h), h->npolys,
                           P<uint32_t>(h->pixedges));
        HIPCHK(h, hipGetLastError());
    }
    // the set's shape bucket: half-octaves of its mean edges per BORDER pixel
    uint64_t *shp = nullptr;
    if ((rc = ensure(h, h->shape_ctr, 16))) return rc;
    shp = P<uint64_t>(h->shape_ctr);
    HIPCHK(h, hipMemsetAsync(shp, 0, 16, h->stream));
    if (h->nstatus > 0) {
        hipLaunchKernelGGL(k_shape_stats, dim3((unsigned)std::min<uint64_t>(blocks_for(h->nstatus), 4096)), dim3(256),
                           0, h->stream, (const uint32_t *)P<uint32_t>(h->pixedges), h->nstatus, shp);
        HIPCHK(h, hipGetLastError());
    }
    uint64_t hs[2] = {0, 0};
    HIPCHK(h, hipMemcpyAsync(hs, shp, 16, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->shape_bucket = hs[1] ? (int)std::floor(2.0 * std::log2(std::max(1.0, (double)hs[0] / (double)hs[1]))) : 0;
    // the bucket's tuned configuration, if the table holds one
    auto it = h->tunings.find(h->shape_bucket);
    if (it != h->tunings.end()) {
        h->tuned_tile = it->second.tile_a;
        h->tuned_chunk = it->second.chunk;
        h->tuned_bpc = it->second.blocks_per_cu;
    }
    return build_chunk_boxes(h, (uint32_t)h->tuned_chunk);
}

// Boxes of the refinement's C-step target chunks, per edge sequence.
// Synchronises the stream; device arrays move, so captured graphs go.
int build_chunk_boxes(farm_handle h, uint32_t C) {
    int rc;
    const uint64_t nseqs = h->nseqs;
    drop_graphs(h);
    h->set_gen++;
    if ((rc = ensure(h, h->chcnt, 4 * (nseqs + 1)))) return rc;
    if ((rc = ensure(h, h->choff, 4 * (nseqs + 1)))) return rc;
    if ((rc = fill_u32(h, h->chcnt.p, nseqs + 1, 0))) return rc;
    if (h->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_count, dim3((unsigned)h->npolys), dim3(256), 0, h->stream, devset(h), h->npolys, C,
                           P<uint32_t>(h->chcnt));
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = exclusive_scan(h, P<uint32_t>(h->chcnt), P<uint32_t>(h->choff), nseqs + 1))) return rc;
    uint32_t nchunks = 0;
    HIPCHK(h, hipMemcpyAsync(&nchunks, P<uint32_t>(h->choff) + nseqs, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if ((rc = ensure(h, h->chbox, 16ull * (nchunks + kRefScreen)))) return rc;
    if (h->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_box, dim3((unsigned)h->npolys), dim3(256), 0, h->stream, devset(h), h->npolys, C,
                           (const uint32_t *)P<uint32_t>(h->choff), P<float4>(h->chbox));
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->chunk = C;
    return FARM_OK;
}

// The intersection's kIxChunk-segment chunk boxes, when the refinement's are
// of another size (else those serve).  Synchronises the stream.
int build_ix_boxes(farm_handle h) {
    int rc;
    if (h->chunk == kIxChunk || h->ix_gen == h->set_gen) return FARM_OK;
    const uint64_t nseqs = h->nseqs;
    if ((rc = ensure(h, h->ixcnt, 4 * (nseqs + 1)))) return rc;
    if ((rc = ensure(h, h->ixoff, 4 * (nseqs + 1)))) return rc;
    if ((rc = fill_u32(h, h->ixcnt.p, nseqs + 1, 0))) return rc;
    if (h->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_count, dim3((unsigned)h->npolys), dim3(256), 0, h->stream, devset(h), h->npolys,
                           kIxChunk, P<uint32_t>(h->ixcnt));
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = exclusive_scan(h, P<uint32_t>(h->ixcnt), P<uint32_t>(h->ixoff), nseqs + 1))) return rc;
    uint32_t nchunks = 0;
    HIPCHK(h, hipMemcpyAsync(&nchunks, P<uint32_t>(h->ixoff) + nseqs, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if ((rc = ensure(h, h->ixbox, 16ull * (nchunks + 1)))) return rc;
    if (h->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_box, dim3((unsigned)h->npolys), dim3(256), 0, h->stream, devset(h), h->npolys,
                           kIxChunk, (const uint32_t *)P<uint32_t>(h->ixoff), P<float4>(h->ixbox));
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ix_gen = h->set_gen;
    return FARM_OK;
}

}  // namespace farm

extern "C" {

enum { OUT_NONE = 0, OUT_DEV = 1, OUT_HOST = 2 };
static int run_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, int out_kind,
                         uint8_t *out, farm_stats *st, uint32_t *pair_tests = nullptr, bool fixed = false);

const char *farm_version(void) { return "farm-mi355x 0.1 (gfx950)"; }

int farm_copy_probe(farm_handle h, uint64_t bytes, int iters, double *gbps) {
    if (!h || iters < 1 || bytes < 16) return fail(h, FARM_EINVAL, "bad copy probe arguments");
    HIPCHK(h, hipSetDevice(h->device));
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return fail(h, FARM_ENOMEM, "copy probe: hipMalloc failed");
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(h, FARM_ENOMEM, "copy probe: hipMalloc failed");
    }
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)h->num_cus * 32);
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, h->stream, (const uint4 *)a, (uint4 *)b, n);
    (void)hipEventRecord(h->ev[0], h->stream);
    for (int k = 0; k < iters; k++)
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, h->stream, (const uint4 *)a, (uint4 *)b, n);
    (void)hipEventRecord(h->ev[1], h->stream);
    hipError_t e = hipStreamSynchronize(h->stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(h, FARM_EDEVICE, "copy probe: %s", hipGetErrorString(e));
    if (gbps) *gbps = 2.0 * n * 16 * iters / (ms / 1000.0) / 1e9;
    return FARM_OK;
}

static bool valid_tuning(const farm_tuning &t) {
    return (t.tile_a == 1 || t.tile_a == 2 || t.tile_a == 4 || t.tile_a == 8) &&
           (t.chunk == 8 || t.chunk == 16 || t.chunk == 32) && t.blocks_per_cu >= 1 && t.blocks_per_cu <= 32;
}

// Make t the default configuration and the entry of the set's bucket.
static int apply_tuning(farm_handle h, const farm_tuning &t) {
    h->tuned_tile = t.tile_a;
    h->tuned_chunk = t.chunk;
    h->tuned_bpc = t.blocks_per_cu;
    farm_tuning e = t;
    e.shape_bucket = h->shape_bucket;
    h->tunings[h->shape_bucket] = e;
    if (h->npolys > 0 && (uint32_t)t.chunk != h->chunk) return build_chunk_boxes(h, (uint32_t)t.chunk);
    return FARM_OK;
}

int farm_autotune_config(farm_handle h, uint64_t begin, uint64_t end, farm_tuning *best) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    HIPCHK(h, hipSetDevice(h->device));
    // time one representative pair window with every tile width and chunk
    // size (the chunk-major order rebuilds the boxes three times), then the
    // workgroups per CU at the fastest pair; each configuration runs twice
    // and its second (warm) time counts
    const int saved_bpc = h->tuned_bpc;
    auto time_cfg = [&](int E, int C, int bpc, double *ms) {
        farm_opts o;
        memset(&o, 0, sizeof o);
        o.mode = FARM_MODE_EXACT_BATCHED;
        o.tile_a = E;
        o.tile_b = C;
        h->tuned_bpc = bpc;
        farm_stats st;
        int rc = run_intersect(h, begin, end, &o, OUT_NONE, nullptr, &st);
        if (!rc) rc = run_intersect(h, begin, end, &o, OUT_NONE, nullptr, &st);
        *ms = st.total_ms;
        return rc;
    };
    farm_tuning b;
    memset(&b, 0, sizeof b);
    b.ms = 1e300;
    int rc = FARM_OK;
    for (int C : {8, 16, 32}) {
        for (int E : {1, 2, 4, 8}) {
            double ms = 0;
            if ((rc = time_cfg(E, C, 8, &ms))) break;
            if (ms < b.ms) {
                b.tile_a = E;
                b.chunk = C;
                b.blocks_per_cu = 8;
                b.ms = ms;
            }
        }
        if (rc) break;
    }
    if (!rc)
        for (int bpc : {4, 16}) {
            double ms = 0;
            if ((rc = time_cfg(b.tile_a, b.chunk, bpc, &ms))) break;
            if (ms < b.ms) {
                b.blocks_per_cu = bpc;
                b.ms = ms;
            }
        }
    h->tuned_bpc = saved_bpc;
    if (rc) return rc;
    if ((rc = apply_tuning(h, b))) return rc;
    b.shape_bucket = h->shape_bucket;
    if (best) *best = b;
    return FARM_OK;
}

int farm_autotune(farm_handle h, uint64_t begin, uint64_t end, int *best_tile) {
    farm_tuning t;
    const int rc = farm_autotune_config(h, begin, end, &t);
    if (!rc && best_tile) *best_tile = t.tile_a;
    return rc;
}

int farm_set_tile(farm_handle h, int tile) {
    if (!h) return FARM_EINVAL;
    if (tile != 1 && tile != 2 && tile != 4 && tile != 8) return fail(h, FARM_EINVAL, "tile must be 1, 2, 4 or 8");
    h->tuned_tile = tile;
    return FARM_OK;
}

int farm_get_tuning(farm_handle h, farm_tuning *out) {
    if (!h || !out) return FARM_EINVAL;
    memset(out, 0, sizeof *out);
    out->tile_a = h->tuned_tile;
    out->chunk = h->tuned_chunk;
    out->blocks_per_cu = h->tuned_bpc;
    out->shape_bucket = h->shape_bucket;
    auto it = h->tunings.find(h->shape_bucket);
    if (it != h->tunings.end() && it->second.tile_a == out->tile_a && it->second.chunk == out->chunk &&
        it->second.blocks_per_cu == out->blocks_per_cu)
        out->ms = it->second.ms;
    return FARM_OK;
}

int farm_set_tuning(farm_handle h, const farm_tuning *t) {
    if (!h || !t) return FARM_EINVAL;
    if (!valid_tuning(*t))
        return fail(h, FARM_EINVAL, "tuning (tile %d, chunk %d, %d per CU): tile 1, 2, 4 or 8, chunk 8, 16 or 32, "
                                    "1..32 workgroups per CU", t->tile_a, t->chunk, t->blocks_per_cu);
    HIPCHK(h, hipSetDevice(h->device));
    return apply_tuning(h, *t);
}

int farm_tuning_save(farm_handle h, const char *path) {
    if (!h || !path) return FARM_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f) return fail(h, FARM_EINVAL, "cannot write %s", path);
    fprintf(f, "farm-tuning 1\n");
    for (const auto &kv : h->tunings)
        fprintf(f, "%d %d %d %d %.6f\n", kv.first, kv.second.tile_a, kv.second.chunk, kv.second.blocks_per_cu,
                kv.second.ms);
    const bool ok = fclose(f) == 0;
    return ok ? FARM_OK : fail(h, FARM_EINVAL, "cannot write %s", path);
}

int farm_tuning_load(farm_handle h, const char *path) {
    if (!h || !path) return FARM_EINVAL;
    FILE *f = fopen(path, "r");
    if (!f) return fail(h, FARM_EINVAL, "cannot read %s", path);
    char line[256];
    int version = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "farm-tuning %d", &version) != 1 || version != 1) {
        fclose(f);
        return fail(h, FARM_EINVAL, "%s: not a farm tuning file", path);
    }
    std::map<int, farm_tuning> got;
    int lineno = 1;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        farm_tuning t;
        memset(&t, 0, sizeof t);
        char tail = 0;
        const int n = sscanf(line, "%d %d %d %d %lf %c", &t.shape_bucket, &t.tile_a, &t.chunk, &t.blocks_per_cu,
                             &t.ms, &tail);
        if (n == EOF || (n == 0 && line[0] == '\n')) continue;
        if (n != 5 || !valid_tuning(t)) {
            fclose(f);
            return fail(h, FARM_EINVAL, "%s:%d: malformed tuning entry", path, lineno);
        }
        got[t.shape_bucket] = t;
    }
    fclose(f);
    for (const auto &kv : got) h->tunings[kv.first] = kv.second;
    auto it = h->tunings.find(h->shape_bucket);
    if (h->npolys > 0 && it != h->tunings.end()) {
        HIPCHK(h, hipSetDevice(h->device));
        return apply_tuning(h, it->second);
    }
    return FARM_OK;
}

int farm_open(int device, farm_handle *out) {
    if (!out) return FARM_EINVAL;
    *out = nullptr;
    farm_handle h = new farm_handle_s();
    h->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        delete h;
        return FARM_EDEVICE;
    }
    if (hipSetDevice(device) != hipSuccess) {
        delete h;
        return FARM_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete h;
        return FARM_EDEVICE;
    }
    h->num_cus = prop.multiProcessorCount;
    // this device's copy of the fixed scratch array (farm_intersect_noalloc):
    // its device address, looked up once per device with the device current
    {
        static std::mutex mu;
        static std::vector<void *> addr;
        std::lock_guard<std::mutex> lock(mu);
        if ((size_t)device >= addr.size()) addr.resize(device + 1, nullptr);
        if (!addr[device]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(farm::scratch)) != hipSuccess || !p) {
                delete h;
                return FARM_EDEVICE;
            }
            addr[device] = p;
        }
        h->scratch = addr[device];
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return FARM_EDEVICE;
    }
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete h;
            return FARM_EDEVICE;
        }
    *out = h;
    return FARM_OK;
}

int farm_close(farm_handle h) {
    if (!h) return FARM_EINVAL;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (int k = 0; k < 2; k++) {
        if (h->stage[k]) (void)hipHostFree(h->stage[k]);
        if (h->stage_ev[k]) (void)hipEventDestroy(h->stage_ev[k]);
    }
    DevBuf *bufs[] = {&h->hdr,    &h->status,  &h->pixoff, &h->seqs,   &h->x,      &h->y,        &h->xyd,      &h->fill,     &h->choff,  &h->chcnt,  &h->chbox,  &h->ixcnt, &h->ixoff, &h->ixbox, &h->voff,
                      &h->vnodes, &h->pixedges, &h->shape_ctr, &h->pa, &h->pb, &h->isx,
                      &h->scan_tmp, &h->pair_tests, &h->mlp_w, &h->mlp_m, &h->mlp_v, &h->mlp_g,
                      &h->mlp_norm, &h->mlp_x, &h->mlp_h1, &h->mlp_h2, &h->mlp_z3, &h->mlp_d3, &h->mlp_dh2,
                      &h->mlp_dh1, &h->mlp_y, &h->mlp_c, &h->mlp_loss, &h->feat, &h->decided, &h->pred,
                      &h->skipbuf, &h->layers, &h->lstatus, &h->areas, &h->wctr, &h->wovf, &h->wslab,
                      &h->wslab_big, &h->wout, &h->wmd, &h->ring2a, &h->icounts, &h->iarea, &h->ipieces, &h->cg_boxa,
                      &h->cg_boxq, &h->cg_cellcnt, &h->cg_celloff, &h->cg_items, &h->cg_cnt, &h->cg_off,
                      &h->cg_keys, &h->cg_keys2, &h->cg_uoff};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FARM_OK;
}

const char *farm_last_error(farm_handle h) { return h ? h->err.c_str() : "null handle"; }

int farm_upload_polygons(farm_handle h, const farm_polyset *ps) {
    if (!h || !ps || ps->npolys < 0) return fail(h, FARM_EINVAL, "bad polygon set");
    HIPCHK(h, hipSetDevice(h->device));
    // validate the layout on the host before any kernel reads it
    for (int64_t i = 0; i < ps->npolys; i++) {
        const farm_raster_hdr &r = ps->hdr[i];
        const uint64_t npix = (uint64_t)r.dimx * (uint64_t)r.dimy;
        if (r.dimx <= 0 || r.dimy <= 0 || r.pix_start + npix > ps->nstatus ||
            r.pixoff_start + npix + 1 > ps->npixoff || r.seq_start + r.nseqs > ps->nseqs ||
            r.vert_start + r.nverts > ps->nverts || r.nverts < 2 || r.vline_start + r.dimx + 2 > ps->nvoff ||
            !(r.step[0] > 0) || !(r.step[1] > 0))
            return fail(h, FARM_EINVAL, "polygon %lld: raster header out of range", (long long)i);
        const uint32_t *po = ps->pixoff + r.pixoff_start;
        for (uint64_t p = 0; p <= npix; p++)
            if (po[p] > r.nseqs || (p < npix && po[p] > po[p + 1]))
                return fail(h, FARM_EINVAL, "polygon %lld: pixel offsets out of range", (long long)i);
        const uint32_t *vo = ps->voff + r.vline_start;
        for (int p = 0; p < r.dimx + 2; p++)
            if (r.vnode_start + vo[p] > ps->nvnodes)
                return fail(h, FARM_EINVAL, "polygon %lld: grid-line offsets out of range", (long long)i);
        const uint32_t *sq = ps->seqs + 2 * r.seq_start;
        for (uint32_t k = 0; k < r.nseqs; k++)
            if ((uint64_t)sq[2 * k] + sq[2 * k + 1] + 1 > r.nverts)
                return fail(h, FARM_EINVAL, "polygon %lld: edge sequence out of range", (long long)i);
    }
    // hierarchy: every polygon's layers in range, square steps
    const bool layered = ps->areas != nullptr;
    uint64_t max_layer_pix = 0, max_border_pix = 0;
    if (layered) {
        for (int64_t i = 0; i < ps->npolys; i++) {
            const farm_raster_hdr &r = ps->hdr[i];
            if (r.flags & FARM_HDR_REJECTED) continue;  // within answers 2 without reading it
            if (r.nlayers < 1 || r.layer_start + (uint64_t)(r.nlayers - 1) > ps->nlayers ||
                (r.nlayers > 1 && !ps->layers) || r.step[0] != r.step[1])
                return fail(h, FARM_EINVAL, "polygon %lld: layer header out of range", (long long)i);
            max_layer_pix = std::max<uint64_t>(max_layer_pix, (uint64_t)r.dimx * r.dimy);
            uint64_t nb = 0;
            for (uint64_t p = 0; p < (uint64_t)r.dimx * r.dimy; p++) {
                const uint8_t c = ps->status[r.pix_start + p];
                nb += (c != 0 && c != FARM_CATEGORY_COUNT - 1);
            }
            max_border_pix = std::max(max_border_pix, nb);
            for (int k = 0; k + 1 < r.nlayers; k++) {
                const farm_layer &L = ps->layers[r.layer_start + k];
                const uint64_t np = (uint64_t)L.dimx * L.dimy;
                if (L.dimx <= 0 || L.dimy <= 0 || !(L.step > 0) || L.status_start + np > ps->nlstatus)
                    return fail(h, FARM_EINVAL, "polygon %lld: layer %d out of range", (long long)i, k);
                max_layer_pix = std::max(max_layer_pix, np);
            }
        }
    }
    const int cc = ps->category_count ? ps->category_count : FARM_CATEGORY_COUNT;
    if (cc < 4 || cc > 256 || (cc & (cc - 1)) || (cc != FARM_CATEGORY_COUNT && !ps->fill))
        return fail(h, FARM_EINVAL, "category count %d: a power of 2 in 4..256, with fill classes unless 16", cc);
    int rc;
    if ((rc = upload(h, h->hdr, ps->hdr, sizeof(farm_raster_hdr) * ps->npolys))) return rc;
    if (cc != FARM_CATEGORY_COUNT && (rc = upload(h, h->fill, ps->fill, ps->nstatus))) return rc;
    h->cc = cc;
    if ((rc = upload(h, h->status, ps->status, ps->nstatus))) return rc;
    if ((rc = upload(h, h->pixoff, ps->pixoff, 4 * ps->npixoff))) return rc;
    if ((rc = upload(h, h->seqs, ps->seqs, 8 * ps->nseqs))) return rc;
    // vertex arrays carry kVertexPad zeroed floats past the end: the
    // refinement kernel reads target vertices in fixed-size chunks
    for (DevBuf *b : {&h->x, &h->y})
        if ((rc = ensure(h, *b, 4 * (ps->nverts + kVertexPad)))) return rc;
    if ((rc = upload(h, h->x, ps->x, 4 * ps->nverts))) return rc;
    if ((rc = upload(h, h->y, ps->y, 4 * ps->nverts))) return rc;
    if ((rc = upload(h, h->voff, ps->voff, 4 * ps->nvoff))) return rc;
    if ((rc = upload(h, h->vnodes, ps->vnodes, 8 * ps->nvnodes))) return rc;
    h->has_layers = false;
    h->areas_ready = layered;
    if (layered) {
        if ((rc = upload(h, h->layers, ps->layers, sizeof(farm_layer) * ps->nlayers))) return rc;
        if ((rc = upload(h, h->lstatus, ps->lstatus, ps->nlstatus))) return rc;
        if ((rc = upload(h, h->areas, ps->areas, 8 * ps->nstatus))) return rc;
        h->has_layers = true;
        h->max_layer_pix = max_layer_pix;
        h->max_border_pix = max_border_pix;
    }
    h->npolys = ps->npolys;
    h->nstatus = ps->nstatus;
    h->nverts = ps->nverts;
    h->nseqs = ps->nseqs;
    h->npixoff = ps->npixoff;
    h->nvoff = ps->nvoff;
    h->nvnodes = ps->nvnodes;
    h->nlayers = layered ? ps->nlayers : 0;
    h->nlstatus = layered ? ps->nlstatus : 0;
    return finish_set(h);
}

int farm_upload_pairs(farm_handle h, const uint32_t *a, const uint32_t *b, uint64_t n) {
    if (!h || (n && (!a || !b))) return fail(h, FARM_EINVAL, "bad pair arrays");
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "upload polygons before pairs");
    for (uint64_t i = 0; i < n; i++)
        if ((int64_t)a[i] >= h->npolys || (int64_t)b[i] >= h->npolys)
            return fail(h, FARM_EINVAL, "pair %llu references polygon outside the set", (unsigned long long)i);
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = upload(h, h->pa, a, 4 * n))) return rc;
    if ((rc = upload(h, h->pb, b, 4 * n))) return rc;
    h->npairs = n;
    h->set_gen++;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

static int run_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, int out_kind,
                         uint8_t *out, farm_stats *st, uint32_t *pair_tests, bool fixed) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    if (begin > end || end > h->npairs) return fail(h, FARM_EINVAL, "pair window [%llu,%llu) outside [0,%llu)",
                                                    (unsigned long long)begin, (unsigned long long)end,
                                                    (unsigned long long)h->npairs);
    const uint64_t n64 = end - begin;
    if (n64 > 0x7fffffffull) return fail(h, FARM_ECAPACITY, "batch of %llu pairs: at most 2^31 - 1 per call",
                                         (unsigned long long)n64);
    const uint32_t n = (uint32_t)n64;
    const int mode = opts ? opts->mode : FARM_MODE_EXACT_BATCHED;
    int E = opts && opts->tile_a ? opts->tile_a : h->tuned_tile;
    if (E != 1 && E != 2 && E != 4 && E != 8) return fail(h, FARM_EINVAL, "tile_a must be 1, 2, 4 or 8");
    // chunk size: the call's, else the tuned one (the built one for the
    // fixed-scratch entry, which never rebuilds the boxes on its own)
    const int C = opts && opts->tile_b ? opts->tile_b : (fixed ? (int)h->chunk : h->tuned_chunk);
    if (C != 8 && C != 16 && C != 32) return fail(h, FARM_EINVAL, "tile_b (chunk size) must be 8, 16 or 32");
    if (mode < FARM_MODE_EXACT_PLAIN || mode > FARM_MODE_APPROX_LEARNED)
        return fail(h, FARM_EINVAL, "mode %d not supported by farm_intersect", mode);
    const bool approx = mode == FARM_MODE_APPROX_HEURISTIC;
    const bool learned = mode == FARM_MODE_APPROX_LEARNED;
    const bool learned_order = opts && opts->order == 1 && mode == FARM_MODE_EXACT_BATCHED;
    if (fixed) {
        if (learned || learned_order)
            return fail(h, FARM_EINVAL, "the learned modes need the MLP's own buffers: use farm_intersect_dev");
        if (n64 > kNaPairs)
            return fail(h, FARM_ECAPACITY, "%llu pairs exceed the fixed scratch's %llu per call",
                        (unsigned long long)n64, (unsigned long long)kNaPairs);
    }
    HIPCHK(h, hipSetDevice(h->device));
    if ((uint32_t)C != h->chunk && mode != FARM_MODE_EXACT_PLAIN && mode != FARM_MODE_APPROX_HEURISTIC) {
        // boxes for another chunk size: rebuilt here, which allocates
        if (fixed)
            return fail(h, FARM_EINVAL, "chunk size %d: the set's boxes are built for %u (farm_set_tuning first)", C,
                        h->chunk);
        int rc = build_chunk_boxes(h, (uint32_t)C);
        if (rc) return rc;
    }
    const DevSet S = devset(h);
    unsigned long long c[kCtrSlots];
    IsxWork W;
    for (int attempt = 0;; attempt++) {
        int rc;
        if (fixed) W = isx_carve(h->scratch, kNaPairs, kNaPixPairs, kNaTiles);
        else if ((rc = isx_reserve(h, n, &W))) return rc;
        uint8_t *outp = out_kind == OUT_DEV ? out : (out_kind == OUT_HOST ? W.bits : nullptr);
        hipLaunchKernelGGL(k_set_desc, dim3(1), dim3(64), 0, h->stream, W.desc, (unsigned long long)begin,
                           (unsigned long long)n, outp, pair_tests);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
        if (mode == FARM_MODE_EXACT_PLAIN) {
            if ((rc = fill_u32(h, W.ctr, 2 * kCtrSlots, 0))) return rc;
            if ((rc = fill_u32(h, W.result, n, 0))) return rc;
            if (pair_tests && (rc = fill_u32(h, pair_tests, n, 0))) return rc;
            HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
            HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
            if (n) {
                hipLaunchKernelGGL(k_plain, dim3(n), dim3(256), 0, h->stream, S,
                                   (const uint32_t *)P<uint32_t>(h->pa) + begin,
                                   (const uint32_t *)P<uint32_t>(h->pb) + begin, n, W.result, W.ctr);
                HIPCHK(h, hipGetLastError());
            }
        } else {
            IsxCfg cfg;
            memset(&cfg, 0, sizeof cfg);
            cfg.approx = approx ? 1 : 0;
            cfg.tile = approx ? 0 : E;
            if (learned || learned_order) {
                // pair features + MLP on the device (timed with the filter)
                if ((rc = farm_mlp_predict_pairs(h, begin, end, nullptr))) return rc;
                if (learned_order) cfg.pred_cost = P<float>(h->pred) + 1;
            }
            if (learned && n) {
                if ((rc = ensure(h, h->skipbuf, n))) return rc;
                const float accept = opts->approx_threshold;
                const float reject = opts->approx_reject < accept ? opts->approx_reject : accept;
                hipLaunchKernelGGL(k_learned_decide, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                                   (const int8_t *)P<int8_t>(h->decided), (const float *)P<float>(h->pred), n, accept,
                                   reject, W.result, P<uint8_t>(h->skipbuf));
                HIPCHK(h, hipGetLastError());
                cfg.skip = P<uint8_t>(h->skipbuf);
            }
            cfg.ws = W.desc;
            cfg.cap_pairs = W.cap_pairs;
            cfg.cap_pp = W.cap_pp;
            cfg.cap_tiles = W.cap_tiles;
            cfg.set_gen = h->set_gen;
            cfg.pa = P<uint32_t>(h->pa);
            cfg.pb = P<uint32_t>(h->pb);
            cfg.grid = h->num_cus * 16;
            cfg.prune = opts && (opts->flags & FARM_FLAG_NO_PRUNE) ? 0 : 1;
            cfg.chunk = approx ? 0 : C;
            cfg.bpc = approx ? 0 : h->tuned_bpc;
            rc = (h->use_graphs && !fixed) ? isx_front_graph(h, S, W, cfg) : isx_front(h, S, W, cfg);
            if (rc) return rc;
            HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
            HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
            if (!approx && (rc = isx_refine(h, S, W, cfg))) return rc;
        }
        HIPCHK(h, hipEventRecord(h->ev[3], h->stream));
        hipLaunchKernelGGL(k_finalize, dim3(h->num_cus * 4), dim3(256), 0, h->stream, (const uint32_t *)W.result,
                           (const IsxDesc *)W.desc, W.ctr);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(c, W.ctr, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (mode == FARM_MODE_EXACT_PLAIN || approx) break;
        // the pixel-pair list and the tile map are bounded buffers: a window
        // that overflows them ran no refinement and is repeated with grown
        // buffers, or fails when it exceeds 32-bit offsets or the limit
        const unsigned long long npp = c[kCtrPixPairs], ntiles = c[kCtrTiles];
        if (npp > h->max_pixel_pairs || npp > 0xffffffffull)
            return fail(h, FARM_ECAPACITY,
                        "%llu BORDER x BORDER pixel pairs in pairs [%llu,%llu) exceed the per-call limit of %llu; "
                        "use smaller batches",
                        npp, (unsigned long long)begin, (unsigned long long)end,
                        (unsigned long long)std::min<uint64_t>(h->max_pixel_pairs, 0xffffffffull));
        if (npp > W.cap_pp) {
            if (fixed)
                return fail(h, FARM_ECAPACITY, "%llu pixel pairs exceed the fixed scratch's %llu per call", npp,
                            (unsigned long long)kNaPixPairs);
            h->isx_want_pp = npp + npp / 4 + 4096;
            h->isx_want_tiles = std::max(h->isx_want_tiles, h->isx_want_pp);
            continue;
        }
        if (ntiles > 0xffffffffull)
            return fail(h, FARM_ECAPACITY, "%llu refinement tiles exceed 32-bit offsets; use smaller batches", ntiles);
        if (ntiles > W.cap_tiles) {
            if (fixed)
                return fail(h, FARM_ECAPACITY, "%llu refinement tiles exceed the fixed scratch's %llu per call",
                            ntiles, (unsigned long long)kNaTiles);
            h->isx_want_tiles = ntiles + ntiles / 4 + 4096;
            continue;
        }
        break;
    }
    if (st) {
        memset(st, 0, sizeof *st);
        st->pairs = n;
        st->filter_bytes = c[kCtrFilterBytes];
        st->in_hits = c[kCtrInHits];
        st->bb_pixel_pairs = c[kCtrPixPairs];
        st->segment_tests = c[kCtrTests];
        st->refine_items = c[kCtrItems];
        st->refine_bytes = c[kCtrRefineBytes];
        st->found = c[kCtrFound];
        st->pruned_tests = c[kCtrPruned];
        float t01 = 0, t23 = 0, t03 = 0;
        HIPCHK(h, hipEventElapsedTime(&t01, h->ev[0], h->ev[1]));
        HIPCHK(h, hipEventElapsedTime(&t23, h->ev[2], h->ev[3]));
        HIPCHK(h, hipEventElapsedTime(&t03, h->ev[0], h->ev[3]));
        st->filter_ms = t01;
        st->refine_ms = t23;
        st->total_ms = t03;
    }
    if (out_kind == OUT_HOST && n) HIPCHK(h, hipMemcpy(out, W.bits, n, hipMemcpyDeviceToHost));
    return FARM_OK;
}

int farm_intersect_dev(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                       farm_stats *st) {
    return run_intersect(h, begin, end, opts, out_dev ? OUT_DEV : OUT_NONE, out_dev, st);
}

int farm_intersect_noalloc(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                           farm_stats *st) {
    return run_intersect(h, begin, end, opts, out_dev ? OUT_DEV : OUT_NONE, out_dev, st, nullptr, true);
}

int farm_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out,
                   farm_stats *st) {
    return run_intersect(h, begin, end, opts, out ? OUT_HOST : OUT_NONE, out, st);
}

int farm_intersect_costs(farm_handle h, uint64_t begin, uint64_t end, uint8_t *bits, uint32_t *tests,
                         farm_stats *st) {
    if (!h) return FARM_EINVAL;
    const uint64_t n = end > begin ? end - begin : 0;
    int rc = ensure(h, h->pair_tests, 4 * (n ? n : 1));
    if (rc) return rc;
    farm_opts o;
    memset(&o, 0, sizeof o);
    o.mode = FARM_MODE_EXACT_BATCHED;
    rc = run_intersect(h, begin, end, &o, bits ? OUT_HOST : OUT_NONE, bits, st, P<uint32_t>(h->pair_tests));
    if (rc == FARM_OK && n && tests) HIPCHK(h, hipMemcpy(tests, h->pair_tests.p, 4 * n, hipMemcpyDeviceToHost));
    return rc;
}

int farm_set_pixel_pair_limit(farm_handle h, uint64_t max_pixel_pairs) {
    if (!h) return FARM_EINVAL;
    h->max_pixel_pairs = max_pixel_pairs ? std::min<uint64_t>(max_pixel_pairs, 0xffffffffull) : 0xffffffffull;
    return FARM_OK;
}

int farm_set_graphs(farm_handle h, int on) {
    if (!h) return FARM_EINVAL;
    h->use_graphs = on != 0;  // captured graphs stay cached for when they are switched on again
    return FARM_OK;
}

}  // extern "C"
