// C ABI of libfarm.so: device memory, uploads and the intersect driver.
// Declarations and their reference counterparts: include/farm_capi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include <string>
#include <vector>

#include "handle.h"

namespace farm {
__global__ void k_filter(DevSet, const uint32_t *, const uint32_t *, uint32_t, int, int, const uint8_t *, uint32_t *,
                         const uint32_t *, PixPair *, uint32_t *, unsigned long long *);
__global__ void k_bucket_hist(DevSet, const uint32_t *, const uint32_t *, const PixPair *, uint32_t, const float *,
                              uint8_t *, uint32_t *);
__global__ void k_bucket_scan(uint32_t *, uint32_t *);
__global__ void k_bucket_scatter(const uint8_t *, uint32_t, uint32_t *, uint32_t *);
template <int E>
__global__ void k_tiles(const PixPair *, const uint32_t *, uint32_t, uint32_t *, unsigned long long *);
template <int E>
__global__ void k_refine(DevSet, const uint32_t *, const uint32_t *, const PixPair *, const uint32_t *,
                         const uint2 *, uint32_t, uint32_t *, uint32_t *, unsigned long long *, uint32_t *);
__global__ void k_tile_map(const uint32_t *, uint32_t, uint32_t, uint2 *);
__global__ void k_plain(DevSet, const uint32_t *, const uint32_t *, uint32_t, uint32_t *, unsigned long long *);
__global__ void k_pixedges(DevSet, int64_t, uint32_t *);
__global__ void k_vdiff(const float *, const float *, uint64_t, float *, float *);
__global__ void k_chunk_count(DevSet, int64_t, uint32_t *);
__global__ void k_chunk_box(DevSet, int64_t, const uint32_t *, float4 *);
__global__ void k_finalize(const uint32_t *, uint32_t, uint8_t *, unsigned long long *);
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

// tiles per ordered pixel pair, their scan, and the persistent refinement
// with E source edges per lane (events ev[2]..ev[3] bracket the refinement)
template <int E>
int launch_tiles_refine(farm_handle h, const DevSet &S, const uint32_t *pa, const uint32_t *pb, uint32_t npp_total,
                        unsigned long long *ctr, uint32_t *pair_tests) {
    hipLaunchKernelGGL(k_tiles<E>, dim3(blocks_for(npp_total)), dim3(256), 0, h->stream,
                       (const PixPair *)P<PixPair>(h->pps), (const uint32_t *)P<uint32_t>(h->order), npp_total,
                       P<uint32_t>(h->tiles), ctr);
    HIPCHK(h, hipGetLastError());
    int rc = exclusive_scan<uint32_t>(h, P<uint32_t>(h->tiles), P<uint32_t>(h->tile_off), npp_total + 1);
    if (rc) return rc;
    // the tile total (32-bit offsets: checked against the 64-bit count first)
    unsigned long long ntiles64 = 0;
    uint32_t ntiles = 0;
    HIPCHK(h, hipMemcpyAsync(&ntiles64, ctr + 7, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&ntiles, P<uint32_t>(h->tile_off) + npp_total, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (ntiles64 > 0xffffffffull)
        return fail(h, FARM_ECAPACITY, "%llu refinement tiles exceed 32-bit offsets; use smaller batches", ntiles64);
    if ((rc = ensure(h, h->tile_map, 8ull * ntiles + 8))) return rc;
    if (npp_total) {
        hipLaunchKernelGGL(k_tile_map, dim3(blocks_for(npp_total)), dim3(256), 0, h->stream,
                           (const uint32_t *)P<uint32_t>(h->tile_off), npp_total, ntiles, P<uint2>(h->tile_map));
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = fill_u32(h, h->cursor.p, 1, 0))) return rc;
    const unsigned grid = (unsigned)(h->num_cus * 8);
    HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
    hipLaunchKernelGGL(k_refine<E>, dim3(grid), dim3(256), 0, h->stream, S, pa, pb,
                       (const PixPair *)P<PixPair>(h->pps), (const uint32_t *)P<uint32_t>(h->order),
                       (const uint2 *)P<uint2>(h->tile_map), ntiles, P<uint32_t>(h->cursor),
                       P<uint32_t>(h->result), ctr, pair_tests);
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

}  // namespace

extern "C" {

static int run_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                         farm_stats *st, uint32_t *pair_tests = nullptr);

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

int farm_autotune(farm_handle h, uint64_t begin, uint64_t end, int *best_tile) {
    if (!h) return FARM_EINVAL;
    // time the refinement of one representative pair window with every tile
    // width, keep the fastest for later calls with opts->tile_a == 0
    const int cand[] = {1, 2, 4, 8};
    double best_ms = 1e300;
    int best = h->tuned_tile;
    for (int E : cand) {
        farm_opts o;
        memset(&o, 0, sizeof o);
        o.mode = FARM_MODE_EXACT_BATCHED;
        o.tile_a = E;
        farm_stats st;
        int rc = run_intersect(h, begin, end, &o, nullptr, &st);
        if (rc) return rc;
        rc = run_intersect(h, begin, end, &o, nullptr, &st);  // second run: warm caches
        if (rc) return rc;
        if (st.total_ms < best_ms) {
            best_ms = st.total_ms;
            best = E;
        }
    }
    h->tuned_tile = best;
    if (best_tile) *best_tile = best;
    return FARM_OK;
}

int farm_set_tile(farm_handle h, int tile) {
    if (!h) return FARM_EINVAL;
    if (tile != 1 && tile != 2 && tile != 4 && tile != 8) return fail(h, FARM_EINVAL, "tile must be 1, 2, 4 or 8");
    h->tuned_tile = tile;
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
    DevBuf *bufs[] = {&h->hdr,    &h->status,  &h->pixoff, &h->seqs,   &h->x,      &h->y,        &h->dx,       &h->dy,     &h->choff,  &h->chcnt,  &h->chbox,  &h->voff,
                      &h->vnodes, &h->pixedges, &h->pa,    &h->pb,     &h->result, &h->npp,      &h->pp_off,
                      &h->pps,    &h->bucket,  &h->order,  &h->tiles,  &h->tile_off, &h->tile_map, &h->ctr,    &h->hist,
                      &h->cursor, &h->scan_tmp, &h->pair_tests, &h->mlp_w, &h->mlp_m, &h->mlp_v, &h->mlp_g,
                      &h->mlp_norm, &h->mlp_x, &h->mlp_h1, &h->mlp_h2, &h->mlp_z3, &h->mlp_d3, &h->mlp_dh2,
                      &h->mlp_dh1, &h->mlp_y, &h->mlp_c, &h->mlp_loss, &h->feat, &h->decided, &h->pred,
                      &h->skipbuf, &h->layers, &h->lstatus, &h->areas, &h->wctr, &h->wovf, &h->wslab,
                      &h->wslab_big, &h->wout, &h->wmd, &h->ring2a, &h->icounts, &h->iarea, &h->cg_boxa,
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
    int rc;
    if ((rc = upload(h, h->hdr, ps->hdr, sizeof(farm_raster_hdr) * ps->npolys))) return rc;
    if ((rc = upload(h, h->status, ps->status, ps->nstatus))) return rc;
    if ((rc = upload(h, h->pixoff, ps->pixoff, 4 * ps->npixoff))) return rc;
    if ((rc = upload(h, h->seqs, ps->seqs, 8 * ps->nseqs))) return rc;
    // vertex arrays carry kVertexPad zeroed floats past the end: the
    // refinement kernel reads target vertices in fixed-size chunks
    for (DevBuf *b : {&h->x, &h->y}) {
        if ((rc = ensure(h, *b, 4 * (ps->nverts + kVertexPad)))) return rc;
        HIPCHK(h, hipMemsetAsync((char *)b->p + 4 * ps->nverts, 0, 4 * kVertexPad, h->stream));
    }
    if ((rc = upload(h, h->x, ps->x, 4 * ps->nverts))) return rc;
    if ((rc = upload(h, h->y, ps->y, 4 * ps->nverts))) return rc;
    // successive vertex differences over the padded arrays (the last one
    // reads no vertex past the padding and is 0)
    for (DevBuf *b : {&h->dx, &h->dy})
        if ((rc = ensure(h, *b, 4 * (ps->nverts + kVertexPad)))) return rc;
    hipLaunchKernelGGL(k_vdiff, dim3(blocks_for(ps->nverts + kVertexPad)), dim3(256), 0, h->stream,
                       P<float>(h->x), P<float>(h->y), ps->nverts + kVertexPad, P<float>(h->dx), P<float>(h->dy));
    HIPCHK(h, hipGetLastError());
    if ((rc = upload(h, h->voff, ps->voff, 4 * ps->nvoff))) return rc;
    if ((rc = upload(h, h->vnodes, ps->vnodes, 8 * ps->nvnodes))) return rc;
    if ((rc = ensure(h, h->pixedges, 4 * ps->nstatus))) return rc;
    h->has_layers = false;
    h->ring_area_ready = false;
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
    if (ps->npolys > 0) {
        hipLaunchKernelGGL(k_pixedges, dim3((unsigned)ps->npolys), dim3(256), 0, h->stream, devset(h), ps->npolys,
                           P<uint32_t>(h->pixedges));
        HIPCHK(h, hipGetLastError());
    }
    // boxes of the refinement's 16-step target chunks, per edge sequence
    if ((rc = ensure(h, h->chcnt, 4 * (ps->nseqs + 1)))) return rc;
    if ((rc = ensure(h, h->choff, 4 * (ps->nseqs + 1)))) return rc;
    if ((rc = fill_u32(h, h->chcnt.p, ps->nseqs + 1, 0))) return rc;
    if (ps->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_count, dim3((unsigned)ps->npolys), dim3(256), 0, h->stream, devset(h),
                           ps->npolys, P<uint32_t>(h->chcnt));
        HIPCHK(h, hipGetLastError());
    }
    if ((rc = exclusive_scan(h, P<uint32_t>(h->chcnt), P<uint32_t>(h->choff), ps->nseqs + 1))) return rc;
    uint32_t nchunks = 0;
    HIPCHK(h, hipMemcpyAsync(&nchunks, P<uint32_t>(h->choff) + ps->nseqs, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if ((rc = ensure(h, h->chbox, 16ull * (nchunks + kRefScreen)))) return rc;
    if (ps->npolys > 0) {
        hipLaunchKernelGGL(k_chunk_box, dim3((unsigned)ps->npolys), dim3(256), 0, h->stream, devset(h), ps->npolys,
                           (const uint32_t *)P<uint32_t>(h->choff), P<float4>(h->chbox));
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
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
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

static int run_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                         farm_stats *st, uint32_t *pair_tests) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    if (begin > end || end > h->npairs) return fail(h, FARM_EINVAL, "pair window [%llu,%llu) outside [0,%llu)",
                                                    (unsigned long long)begin, (unsigned long long)end,
                                                    (unsigned long long)h->npairs);
    const uint64_t n64 = end - begin;
    if (n64 > 0xffffffffull) return fail(h, FARM_ECAPACITY, "batch larger than 2^32 pairs");
    const uint32_t n = (uint32_t)n64;
    const int mode = opts ? opts->mode : FARM_MODE_EXACT_BATCHED;
    int E = opts && opts->tile_a ? opts->tile_a : h->tuned_tile;
    if (E != 1 && E != 2 && E != 4 && E != 8) return fail(h, FARM_EINVAL, "tile_a must be 1, 2, 4 or 8");
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    const uint32_t *pa = P<uint32_t>(h->pa) + begin;
    const uint32_t *pb = P<uint32_t>(h->pb) + begin;
    if ((rc = ensure(h, h->result, 4ull * n + 4))) return rc;
    if ((rc = ensure(h, h->ctr, 8 * 16))) return rc;
    if ((rc = fill_u32(h, h->ctr.p, 32, 0))) return rc;
    DevSet S = devset(h);
    unsigned long long *ctr = P<unsigned long long>(h->ctr);
    HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    float filter_ms = 0, refine_ms = 0;
    if (n == 0) {
        // nothing to do
    } else if (mode == FARM_MODE_EXACT_PLAIN) {
        if ((rc = fill_u32(h, h->result.p, n, 0))) return rc;
        HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
        hipLaunchKernelGGL(k_plain, dim3(n), dim3(256), 0, h->stream, S, pa, pb, n, P<uint32_t>(h->result), ctr);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
    } else if (mode == FARM_MODE_EXACT_BATCHED || mode == FARM_MODE_APPROX_HEURISTIC ||
               mode == FARM_MODE_APPROX_LEARNED) {
        const int approx = mode == FARM_MODE_APPROX_HEURISTIC;
        const bool learned = mode == FARM_MODE_APPROX_LEARNED;
        const bool learned_order = opts && opts->order == 1 && !approx;
        const uint8_t *skip = nullptr;
        const float *pred_cost = nullptr;
        if (learned || learned_order) {
            // pair features + MLP on the device (counted in the filter phase)
            if ((rc = farm_mlp_predict_pairs(h, begin, end, nullptr))) return rc;
            pred_cost = P<float>(h->pred) + 1;
        }
        if (learned) {
            if ((rc = ensure(h, h->skipbuf, n))) return rc;
            const float accept = opts->approx_threshold;
            const float reject = opts->approx_reject < accept ? opts->approx_reject : accept;
            hipLaunchKernelGGL(k_learned_decide, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                               (const int8_t *)P<int8_t>(h->decided), (const float *)P<float>(h->pred), n, accept,
                               reject, P<uint32_t>(h->result), P<uint8_t>(h->skipbuf));
            HIPCHK(h, hipGetLastError());
            skip = P<uint8_t>(h->skipbuf);
        }
        if ((rc = ensure(h, h->npp, 4ull * (n + 1)))) return rc;
        if ((rc = ensure(h, h->pp_off, 4ull * (n + 1)))) return rc;
        if ((rc = fill_u32(h, h->npp.p, n + 1, 0))) return rc;
        if (pair_tests && (rc = fill_u32(h, pair_tests, n, 0))) return rc;
        hipLaunchKernelGGL(k_filter, dim3(blocks_for(n)), dim3(256), 0, h->stream, S, pa, pb, n, 0, approx, skip,
                           P<uint32_t>(h->npp), (const uint32_t *)nullptr, (PixPair *)nullptr,
                           P<uint32_t>(h->result), ctr);
        HIPCHK(h, hipGetLastError());
        uint32_t npp_total = 0;
        if (!approx) {
            if ((rc = exclusive_scan<uint32_t>(h, P<uint32_t>(h->npp), P<uint32_t>(h->pp_off), n + 1))) return rc;
            HIPCHK(h, hipMemcpyAsync(&npp_total, P<uint32_t>(h->pp_off) + n, 4, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
        }
        HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
        if (npp_total > 0) {
            if ((rc = ensure(h, h->pps, sizeof(PixPair) * (size_t)npp_total))) return rc;
            if ((rc = ensure(h, h->bucket, npp_total))) return rc;
            if ((rc = ensure(h, h->order, 4ull * npp_total))) return rc;
            if ((rc = ensure(h, h->tiles, 4ull * (npp_total + 1)))) return rc;
            if ((rc = ensure(h, h->tile_off, 4ull * (npp_total + 1)))) return rc;
            if ((rc = ensure(h, h->hist, 4 * 64))) return rc;
            if ((rc = ensure(h, h->cursor, 4 * 64))) return rc;
            hipLaunchKernelGGL(k_filter, dim3(blocks_for(n)), dim3(256), 0, h->stream, S, pa, pb, n, 1, 0, skip,
                               P<uint32_t>(h->npp), (const uint32_t *)P<uint32_t>(h->pp_off), P<PixPair>(h->pps),
                               P<uint32_t>(h->result), ctr);
            HIPCHK(h, hipGetLastError());
            if ((rc = fill_u32(h, h->hist.p, 64, 0))) return rc;
            hipLaunchKernelGGL(k_bucket_hist, dim3(blocks_for(npp_total)), dim3(256), 0, h->stream, S, pa, pb,
                               (const PixPair *)P<PixPair>(h->pps), npp_total,
                               learned_order ? pred_cost : (const float *)nullptr, P<uint8_t>(h->bucket),
                               P<uint32_t>(h->hist));
            HIPCHK(h, hipGetLastError());
            hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(64), 0, h->stream, P<uint32_t>(h->hist),
                               P<uint32_t>(h->cursor));
            HIPCHK(h, hipGetLastError());
            hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks_for(npp_total)), dim3(256), 0, h->stream,
                               (const uint8_t *)P<uint8_t>(h->bucket), npp_total, P<uint32_t>(h->cursor),
                               P<uint32_t>(h->order));
            HIPCHK(h, hipGetLastError());
            if ((rc = fill_u32(h, P<uint32_t>(h->tiles) + npp_total, 1, 0))) return rc;
            switch (E) {
                case 1: rc = launch_tiles_refine<1>(h, S, pa, pb, npp_total, ctr, pair_tests); break;
                case 2: rc = launch_tiles_refine<2>(h, S, pa, pb, npp_total, ctr, pair_tests); break;
                case 4: rc = launch_tiles_refine<4>(h, S, pa, pb, npp_total, ctr, pair_tests); break;
                default: rc = launch_tiles_refine<8>(h, S, pa, pb, npp_total, ctr, pair_tests); break;
            }
            if (rc) return rc;
        } else {
            HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
        }
    } else {
        return fail(h, FARM_EINVAL, "mode %d not supported by farm_intersect", mode);
    }
    if (n) {
        HIPCHK(h, hipEventRecord(h->ev[3], h->stream));
        hipLaunchKernelGGL(k_finalize, dim3(blocks_for(n)), dim3(256), 0, h->stream,
                           (const uint32_t *)P<uint32_t>(h->result), n, out_dev, ctr);
        HIPCHK(h, hipGetLastError());
    }
    unsigned long long c[16];
    HIPCHK(h, hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (c[7] > 0xffffffffull)
        return fail(h, FARM_ECAPACITY, "%llu refinement tiles exceed 32-bit offsets; use smaller batches", c[7]);
    if (st) {
        memset(st, 0, sizeof *st);
        st->pairs = n;
        st->filter_bytes = c[0];
        st->in_hits = c[1];
        st->segment_tests = c[2];
        st->refine_items = c[3];
        st->refine_bytes = c[4];
        st->found = c[5];
        st->pruned_tests = c[8];
        if (n) {
            float t01 = 0, t23 = 0, t03 = 0;
            HIPCHK(h, hipEventElapsedTime(&t01, h->ev[0], h->ev[1]));
            HIPCHK(h, hipEventElapsedTime(&t23, h->ev[2], h->ev[3]));
            HIPCHK(h, hipEventElapsedTime(&t03, h->ev[0], h->ev[3]));
            filter_ms = t01;
            refine_ms = t23;
            st->filter_ms = filter_ms;
            st->refine_ms = refine_ms;
            st->total_ms = t03;
        }
    }
    return FARM_OK;
}

int farm_intersect_dev(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                       farm_stats *st) {
    return run_intersect(h, begin, end, opts, out_dev, st);
}

int farm_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out,
                   farm_stats *st) {
    if (!h) return FARM_EINVAL;
    if (!out) return run_intersect(h, begin, end, opts, nullptr, st);
    const uint64_t n = end > begin ? end - begin : 0;
    DevBuf tmp;
    int rc = ensure(h, tmp, n ? n : 1);
    if (rc) return rc;
    rc = run_intersect(h, begin, end, opts, (uint8_t *)tmp.p, st);
    if (rc == FARM_OK && n) {
        hipError_t e = hipMemcpy(out, tmp.p, n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(h, FARM_EDEVICE, "copy of results failed: %s", hipGetErrorString(e));
    }
    (void)hipFree(tmp.p);
    return rc;
}

int farm_intersect_costs(farm_handle h, uint64_t begin, uint64_t end, uint8_t *bits, uint32_t *tests,
                         farm_stats *st) {
    if (!h) return FARM_EINVAL;
    const uint64_t n = end > begin ? end - begin : 0;
    int rc = ensure(h, h->pair_tests, 4 * (n ? n : 1));
    if (rc) return rc;
    DevBuf tmp;
    if ((rc = ensure(h, tmp, n ? n : 1))) return rc;
    farm_opts o;
    memset(&o, 0, sizeof o);
    o.mode = FARM_MODE_EXACT_BATCHED;
    rc = run_intersect(h, begin, end, &o, (uint8_t *)tmp.p, st, P<uint32_t>(h->pair_tests));
    if (rc == FARM_OK && n) {
        hipError_t e = hipMemcpy(bits, tmp.p, n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(tests, h->pair_tests.p, 4 * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(h, FARM_EDEVICE, "copy of results failed: %s", hipGetErrorString(e));
    }
    (void)hipFree(tmp.p);
    return rc;
}

}  // extern "C"
