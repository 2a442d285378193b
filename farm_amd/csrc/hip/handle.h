// Internal state of a libfarm handle (one device, one stream) and the host
// helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <map>
#include <vector>

#include "device_layout.h"

namespace farm {

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v);

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

constexpr uint64_t kVertexPad = 64;  // >= the refinement chunk of target vertices

// MLP 80 -> 128 -> 128 -> 2 (intersect logit, log refinement cost)
constexpr int kNFeat = 80;
constexpr int kHidden = 128;
constexpr int kOut = 2;
constexpr int kW1 = kNFeat * kHidden, kB1 = kHidden, kW2 = kHidden * kHidden, kB2 = kHidden, kW3 = kHidden * kOut,
              kB3 = kOut;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kW1, kOffW2 = kOffB1 + kB1, kOffB2 = kOffW2 + kW2, kOffW3 = kOffB2 + kB2,
              kOffB3 = kOffW3 + kW3, kNParams = kOffB3 + kB3;

}  // namespace farm

struct farm_handle_s {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::string err;
    // polygon set
    int64_t npolys = 0;
    farm::DevBuf fill;    // -w fill classes (when the set has them)
    int cc = FARM_CATEGORY_COUNT;
    farm::DevBuf hdr, status, pixoff, seqs, x, y, xyd, choff, chcnt, chbox, voff, vnodes, pixedges, shape_ctr;
    // 16-segment chunk boxes for the intersection's chunk skip when the
    // refinement's boxes are of another size (built for set generation ix_gen)
    farm::DevBuf ixcnt, ixoff, ixbox;
    uint64_t ix_gen = ~0ull;
    uint64_t nstatus = 0;
    // array lengths of the set (farm_polyset's counts)
    uint64_t nverts = 0, nseqs = 0, npixoff = 0, nvoff = 0, nvnodes = 0, nlayers = 0, nlstatus = 0;
    // pairs
    uint64_t npairs = 0;
    farm::DevBuf pa, pb;
    farm::DevBuf scan_tmp;  // hipcub temporaries (uploads, candidate pairs)
    // per-call scratch of the intersect driver: one workspace (isect_work.h),
    // grown when a call's pixel pairs or tiles exceed it
    farm::DevBuf isx;
    uint64_t isx_cap_pairs = 0, isx_cap_pp = 0, isx_cap_tiles = 0;
    uint64_t isx_want_pp = 0, isx_want_tiles = 0;  // capacities a repeated call asked for
    uint64_t max_pixel_pairs = 0xffffffffull;     // per-call limit (farm_set_pixel_pair_limit)
    bool use_graphs = true;                        // capture the batch launches in a hipGraph
    uint64_t set_gen = 0;                          // bumped by every upload that moves device arrays
    struct IsxGraph {
        unsigned char key[128];
        hipGraphExec_t exec;
    };
    std::vector<IsxGraph> graphs;
    void *scratch = nullptr;  // this device's fixed scratch array (farm_intersect_noalloc)
    farm::DevBuf pair_tests;  // per-pair executed segment tests (cost labels)
    // refinement launch configuration used when a call names none: source
    // edges per lane, target segments per chunk box and workgroups per CU
    // (farm_autotune_config / farm_set_tuning); `chunk` is the chunk size the
    // set's boxes are built for, and `tunings` holds the autotuner's choice
    // per shape bucket (farm_tuning_save / farm_tuning_load persist it)
    int tuned_tile = 4;
    int tuned_chunk = 16;
    int tuned_bpc = 8;
    uint32_t chunk = 16;
    int shape_bucket = 0;
    std::map<int, farm_tuning> tunings;
    // learned part
    bool mlp_ready = false;
    farm::DevBuf mlp_w, mlp_m, mlp_v, mlp_g;  // f32 master params, Adam moments, grads
    farm::DevBuf mlp_norm;                    // per feature: mean, 1/std (2 * kNFeat)
    farm::DevBuf mlp_x, mlp_h1, mlp_h2, mlp_z3, mlp_d3, mlp_dh2, mlp_dh1, mlp_y, mlp_c, mlp_loss;
    farm::DevBuf feat, decided, pred;  // per-pair features / decided flags / predictions
    farm::DevBuf skipbuf;              // pairs decided before the exact filter
    uint64_t adam_t = 0;
    // within (-h sets)
    bool has_layers = false;
    bool within_consts_ready = false;
    uint64_t max_layer_pix = 0;   // largest layer (pixels) of the set
    uint64_t max_border_pix = 0;  // most BORDER pixels of a finest layer
    farm::DevBuf layers, lstatus, areas;
    bool areas_ready = false;  // areas hold the set's per-pixel clipped areas
    farm::DevBuf wctr, wovf, wslab, wslab_big, wout, wmd;
    // intersection
    bool ring_area_ready = false;
    farm::DevBuf ring2a, icounts, iarea, ipieces;
    // candidate generation
    farm::DevBuf cg_boxa, cg_boxq, cg_cellcnt, cg_celloff, cg_items, cg_uoff, cg_cnt, cg_off, cg_keys, cg_keys2;
    // pinned staging of the streamed vertex upload (two buffers, their DMA events)
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {};
    uint64_t stage_verts = 1ull << 22;  // vertices per staging buffer (8 B each)
};

namespace farm {

inline int fail(farm_handle h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return farm::fail((h), FARM_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                              __FILE__, __LINE__);                                                         \
    } while (0)

inline int ensure(farm_handle h, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return FARM_OK;
    if (b.p) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        HIPCHK(h, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(h, FARM_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    b.cap = bytes;
    return FARM_OK;
}

inline int upload(farm_handle h, DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(h, b, bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(h, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream));
    return FARM_OK;
}

template <class T>
T *P(DevBuf &b) {
    return reinterpret_cast<T *>(b.p);
}

inline DevSet devset(farm_handle h) {
    DevSet S;
    S.hdr = P<farm_raster_hdr>(h->hdr);
    S.status = P<uint8_t>(h->status);
    S.pixoff = P<uint32_t>(h->pixoff);
    S.seqs = P<uint2>(h->seqs);
    S.x = P<float>(h->x);
    S.y = P<float>(h->y);
    S.xyd = P<float4>(h->xyd);
    S.fill = h->cc != FARM_CATEGORY_COUNT ? P<uint8_t>(h->fill) : P<uint8_t>(h->status);
    S.cc = h->cc;
    S.choff = P<uint32_t>(h->choff);
    S.chbox = P<float4>(h->chbox);
    if (h->chunk == kIxChunk) {
        S.ixoff = S.choff;
        S.ixbox = S.chbox;
    } else if (h->ix_gen == h->set_gen) {
        S.ixoff = P<uint32_t>(h->ixoff);
        S.ixbox = P<float4>(h->ixbox);
    } else {
        S.ixoff = nullptr;
        S.ixbox = nullptr;
    }
    S.pixedges = P<uint32_t>(h->pixedges);
    S.voff = P<uint32_t>(h->voff);
    S.vnodes = P<double>(h->vnodes);
    S.layers = h->has_layers ? P<farm_layer>(h->layers) : nullptr;
    S.lstatus = h->has_layers ? P<uint8_t>(h->lstatus) : nullptr;
    S.areas = h->has_layers ? P<double>(h->areas) : nullptr;
    return S;
}

inline unsigned blocks_for(uint64_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

inline int fill_u32(farm_handle h, void *p, uint64_t n, uint32_t v) {
    if (!n) return FARM_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(256), 0, h->stream, (uint32_t *)p, n, v);
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

// The vertex padding, step differences, per-pixel edge counts and chunk
// boxes of the handle's set (after its arrays are in place).
int finish_set(farm_handle h);

inline void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

}  // namespace farm
