// Exact and approximate intersect kernels for gfx950 (wave64).
//
// Semantics: Ideal::intersect of the reference CPU path (Ideal.cpp:1329-1422):
//   TRUE iff some non-OUT source pixel pa in the source window of the target
//   MBB and some non-OUT target pixel pb in the target window of pa's box have
//   an IN pixel on either side, or some BORDER x BORDER pair (pa, pb) has an
//   edge of one of pa's edge sequences crossing an edge of one of pb's
//   (segment_intersect, geometry_computation.h:215-237, binary32).
// The answer is an OR, so any evaluation order (and any early exit) returns
// the reference's bit.
#include <type_traits>

#include "device_layout.h"
#include "isect_work.h"
#include "wave.h"

namespace farm {

// ------------------------------------------------------------------------
// Raster filter: one lane per candidate pair (grid-stride over the window).
// pass 0: count BORDER x BORDER pixel pairs, decide IN hits (and the
//         reference -a heuristic when approx != 0); the 64-bit total of the
//         pixel pairs goes to ctr[kCtrPixPairs]
// pass 1: emit the pixel pairs at the offsets of an exclusive scan; it does
//         nothing when that total exceeds cap_pp (the buffer's capacity: the
//         host grows it or fails the call, and no offset has wrapped).
// A lane stops walking its pair's windows at the first IN overlap (or, in
// approximate mode, the first decisive BORDER pair): the pair is decided and
// its count is not used.  An undecided pair's windows are walked in full.
// skip (optional): pairs already decided by the caller (their result is left
// untouched and they emit no pixel pairs).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_filter(DevSet S, const uint32_t *__restrict__ pa_all,
                                                 const uint32_t *__restrict__ pb_all,
                                                 const IsxDesc *__restrict__ desc, int pass, int approx,
                                                 const uint8_t *__restrict__ skip, uint32_t *__restrict__ npp,
                                                 const uint32_t *__restrict__ pp_off, PixPair *__restrict__ pps,
                                                 uint64_t cap_pp, uint32_t *__restrict__ result,
                                                 unsigned long long *__restrict__ ctr) {
    if (pass == 1 && ctr[kCtrPixPairs] > cap_pp) return;
    const uint64_t n = desc->n;
    const uint32_t *pa_ids = pa_all + desc->begin;
    const uint32_t *pb_ids = pb_all + desc->begin;
    uint32_t *tests = desc->pair_tests;
    unsigned long long bytes = 0, inhit = 0, bb = 0;
    for (uint64_t kk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < n;
         kk += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)kk;
        if (pass == 0 && tests) tests[k] = 0;
        if (skip && skip[k]) {
            if (pass == 0) npp[k] = 0;
            continue;
        }
        if (pass == 1 && npp[k] == 0) continue;  // decided by the count pass or no BORDER pair
        const uint32_t ia = pa_ids[k], ib = pb_ids[k];
        const farm_raster_hdr A = S.hdr[ia];
        const farm_raster_hdr B = S.hdr[ib];
        const uint8_t *sa = S.status + A.pix_start;
        const uint8_t *sb = S.status + B.pix_start;
        const uint32_t *ea = S.pixedges + A.pix_start;
        const uint32_t *pofa = S.pixoff + A.pixoff_start;
        // the reference's -a (its GPU filter) takes the source window over
        // the target MBB without the 1e-6 shrink (query_intersect.cu:51-54);
        // Ideal::intersect shrinks it (MyRaster.cpp:332-337)
        const Window wa = approx ? window_of(A, B.mbr[0], B.mbr[1], B.mbr[2], B.mbr[3], 0.0)
                                 : window_of(A, B.mbr[0], B.mbr[1], B.mbr[2], B.mbr[3]);
        const double area_a = A.step[0] * A.step[1], area_b = B.step[0] * B.step[1];
        const double amax = fmax(area_a, area_b);
        bool in_hit = false, heur = false;
        uint32_t cnt = 0;
        const uint32_t base = pass == 1 ? pp_off[k] : 0u;
        bytes += 8 + 2 * 56;
        for (int x = wa.x0; x <= wa.x1 && !(in_hit || heur); x++) {
            for (int y = wa.y0; y <= wa.y1 && !(in_hit || heur); y++) {
                const int pa = y * A.dimx + x;
                const uint8_t fa = sa[pa];
                bytes += 1;
                const int st_a = show(fa);
                if (st_a == 0) continue;
                double lx, ly, hx, hy;
                pixel_box(A, x, y, lx, ly, hx, hy);
                const Window wb = window_of(B, lx, ly, hx, hy);
                for (int u = wb.x0; u <= wb.x1 && !(in_hit || heur); u++) {
                    for (int v = wb.y0; v <= wb.y1 && !(in_hit || heur); v++) {
                        const int pb = v * B.dimx + u;
                        const uint8_t fb = sb[pb];
                        bytes += 1;
                        const int st_b = show(fb);
                        if (st_b == 0) continue;
                        if (st_a == 2 || st_b == 2) {
                            in_hit = true;
                            continue;
                        }
                        if (approx) {
                            // fill classes at the set's category count (-w)
                            const double prob = (apx_area_cc(S.fill[A.pix_start + pa], area_a, S.cc) +
                                                 apx_area_cc(S.fill[B.pix_start + pb], area_b, S.cc)) /
                                                amax;
                            if (prob > 1.0) heur = true;
                            continue;
                        }
                        if (pass == 1) {
                            // edge slots of pa: sum of (len + 1) over its sequences
                            const uint32_t nseq = pofa[pa + 1] - pofa[pa];
                            pps[base + cnt] = PixPair{k, (uint32_t)pa, (uint32_t)pb, ea[pa] + nseq};
                        }
                        cnt++;
                    }
                }
            }
        }
        if (pass == 0) {
            const bool decided = in_hit || heur;
            npp[k] = decided ? 0u : cnt;
            result[k] = decided ? 1u : 0u;
            inhit += in_hit ? 1 : 0;
            bb += decided ? 0u : cnt;
        }
    }
    if (pass == 0) {
        // wave-level reduction of the counters, one atomic per wave
        for (int o = 32; o > 0; o >>= 1) {
            bytes += __shfl_xor(bytes, o);
            inhit += __shfl_xor(inhit, o);
            bb += __shfl_xor(bb, o);
        }
        if ((threadIdx.x & 63) == 0) {
            if (bytes) atomicAdd(&ctr[kCtrFilterBytes], bytes);
            if (inhit) atomicAdd(&ctr[kCtrInHits], inhit);
            if (bb) atomicAdd(&ctr[kCtrPixPairs], bb);
        }
    }
}

// ------------------------------------------------------------------------
// Exclusive scan of n u32 values, n = bounded_count(nsrc, cap): out[0..n)
// gets the prefix sums and out[n] the (32-bit) total.  Three kernels:
// kScanParts blocks sum one contiguous segment each, one block scans the
// segment sums, and every block rescans its segment from its offset.
// ------------------------------------------------------------------------
__device__ __forceinline__ void scan_segment(uint64_t n, uint32_t b, uint64_t &s0, uint64_t &s1) {
    const uint64_t seg = (n + kScanParts - 1) / kScanParts;
    s0 = min(n, (uint64_t)b * seg);
    s1 = min(n, s0 + seg);
}

// block-wide sum (256 lanes) delivered to every lane
__device__ __forceinline__ uint32_t block_sum256(uint32_t v, uint32_t *lds4) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    return lds4[0] + lds4[1] + lds4[2] + lds4[3];
}

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t *__restrict__ in,
                                                    const unsigned long long *__restrict__ nsrc, uint64_t cap,
                                                    uint32_t *__restrict__ part) {
    __shared__ uint32_t lds[4];
    uint64_t s0, s1;
    scan_segment(bounded_count(nsrc, cap), blockIdx.x, s0, s1);
    uint32_t s = 0;
    for (uint64_t i = s0 + threadIdx.x; i < s1; i += 256) s += in[i];
    s = block_sum256(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one block of kScanParts lanes
__global__ __launch_bounds__(kScanParts) void k_scan_parts(uint32_t *__restrict__ part,
                                                            const unsigned long long *__restrict__ nsrc,
                                                            uint64_t cap, uint32_t *__restrict__ out) {
    __shared__ uint32_t wsum[kScanParts / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t v = part[t];
    uint32_t tot;
    const uint32_t ex = wave_excl_scan(v, lane, &tot);
    if (lane == 0) wsum[w] = tot;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int i = 0; i < kScanParts / 64; i++) {
        before += i < w ? wsum[i] : 0u;
        all += wsum[i];
    }
    part[t] = before + ex;
    if (t == 0) {
        part[kScanParts] = all;
        out[bounded_count(nsrc, cap)] = all;
    }
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t *__restrict__ in,
                                                     const unsigned long long *__restrict__ nsrc, uint64_t cap,
                                                     const uint32_t *__restrict__ part, uint32_t *__restrict__ out) {
    __shared__ uint32_t wsum[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint64_t s0, s1;
    scan_segment(bounded_count(nsrc, cap), blockIdx.x, s0, s1);
    uint32_t carry = part[blockIdx.x];
    for (uint64_t base = s0; base < s1; base += 1024) {
        // four consecutive values per lane
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t i = base + 4 * (uint64_t)t + j;
            v[j] = i < s1 ? in[i] : 0u;
            sum += v[j];
        }
        uint32_t wtot;
        const uint32_t ex = wave_excl_scan(sum, lane, &wtot);
        __syncthreads();
        if (lane == 0) wsum[w] = wtot;
        __syncthreads();
        uint32_t before = 0;
        for (int i = 0; i < w; i++) before += wsum[i];
        uint32_t run = carry + before + ex;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t i = base + 4 * (uint64_t)t + j;
            if (i < s1) out[i] = run;
            run += v[j];
        }
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    }
}

__global__ void k_set_desc(IsxDesc *d, unsigned long long begin, unsigned long long n, uint8_t *out,
                           uint32_t *pair_tests) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        d->begin = begin;
        d->n = n;
        d->out = out;
        d->pair_tests = pair_tests;
    }
}

// ------------------------------------------------------------------------
// Cost order: counting sort of the pixel pairs by floor(log2(cost)),
// heaviest bucket first.  cost = edge slots of pa x edges of pb.
// ------------------------------------------------------------------------
constexpr int kBuckets = 40;

__device__ __forceinline__ int cost_bucket(uint64_t cost) {
    int b = cost ? 63 - __builtin_clzll(cost) : 0;
    return min(b, kBuckets - 1);
}

__global__ __launch_bounds__(256) void k_bucket_hist(DevSet S, const uint32_t *__restrict__ pb_all,
                                                      const IsxDesc *__restrict__ desc,
                                                      const PixPair *__restrict__ pps,
                                                      const unsigned long long *__restrict__ ctr, uint64_t cap_pp,
                                                      const float *__restrict__ pred_cost,
                                                      uint8_t *__restrict__ bucket, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t n = bounded_count(&ctr[kCtrPixPairs], cap_pp);
    const uint32_t *pb_ids = pb_all + desc->begin;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const PixPair p = pps[i];
        int b;
        if (pred_cost) {
            // learned order: the MLP cost head predicts log2(1 + tests) of
            // the pair; predictions are (logit, cost) pairs, hence stride 2
            const float c = fmaxf(pred_cost[2 * (size_t)p.pair], 0.f);
            b = min((int)c, kBuckets - 1);
        } else {
            const farm_raster_hdr B = S.hdr[pb_ids[p.pair]];
            const uint64_t eb = S.pixedges[B.pix_start + p.pb];
            b = cost_bucket((uint64_t)p.slots * eb);
        }
        bucket[i] = (uint8_t)b;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kBuckets; j += blockDim.x)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

// one block: descending-bucket exclusive scan of the histogram
__global__ void k_bucket_scan(uint32_t *__restrict__ hist, uint32_t *__restrict__ cursor) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t acc = 0;
        for (int b = kBuckets - 1; b >= 0; b--) {
            cursor[b] = acc;
            acc += hist[b];
        }
    }
}

// positions are reserved per block round and bucket (one global atomic per
// non-empty bucket per round), then handed out from LDS
__global__ __launch_bounds__(256) void k_bucket_scatter(const uint8_t *__restrict__ bucket,
                                                         const unsigned long long *__restrict__ ctr, uint64_t cap_pp,
                                                         uint32_t *__restrict__ cursor, uint32_t *__restrict__ order) {
    __shared__ uint32_t cnt[kBuckets], base[kBuckets];
    const uint64_t n = bounded_count(&ctr[kCtrPixPairs], cap_pp);
    for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x; b0 < n; b0 += (uint64_t)gridDim.x * blockDim.x) {
        for (int j = threadIdx.x; j < kBuckets; j += blockDim.x) cnt[j] = 0;
        __syncthreads();
        const uint64_t i = b0 + threadIdx.x;
        int b = 0;
        uint32_t local = 0;
        if (i < n) {
            b = bucket[i];
            local = atomicAdd(&cnt[b], 1u);
        }
        __syncthreads();
        for (int j = threadIdx.x; j < kBuckets; j += blockDim.x)
            if (cnt[j]) base[j] = atomicAdd(&cursor[j], cnt[j]);
        __syncthreads();
        if (i < n) order[base[b] + local] = (uint32_t)i;
        __syncthreads();
    }
}

// tiles per ordered pixel pair (written at position i, scanned afterwards);
// the 64-bit total goes to ctr[kCtrTiles]
template <int E>
__global__ __launch_bounds__(256) void k_tiles(const PixPair *__restrict__ pps, const uint32_t *__restrict__ order,
                                                unsigned long long *__restrict__ ctr, uint64_t cap_pp,
                                                uint32_t *__restrict__ tiles) {
    const uint64_t n = bounded_count(&ctr[kCtrPixPairs], cap_pp);
    unsigned long long t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t slots = pps[order[i]].slots;
        const uint32_t per = 64u * E;
        const uint32_t ti = slots > 1 ? (slots - 1 + per - 1) / per : 0u;
        tiles[i] = ti;
        t += ti;
    }
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(&ctr[kCtrTiles], t);
}

// The descriptor of every tile (its ordered pixel pair, the tile within it,
// the two pixels' sequence ranges and polygons), from the scanned counts;
// nothing is written when the tiles exceed the map's capacity
__global__ __launch_bounds__(256) void k_tile_map(DevSet S, const uint32_t *__restrict__ pa_all,
                                                  const uint32_t *__restrict__ pb_all,
                                                  const IsxDesc *__restrict__ desc, const PixPair *__restrict__ pps,
                                                  const uint32_t *__restrict__ order,
                                                  const uint32_t *__restrict__ tile_off,
                                                  const unsigned long long *__restrict__ ctr, uint64_t cap_pp,
                                                  uint64_t cap_tiles, TileDesc *__restrict__ map) {
    const uint64_t npp = bounded_count(&ctr[kCtrPixPairs], cap_pp);
    if (ctr[kCtrTiles] > cap_tiles) return;
    const uint32_t *pa_ids = pa_all + desc->begin;
    const uint32_t *pb_ids = pb_all + desc->begin;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < npp; q += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t0 = tile_off[q], t1 = tile_off[q + 1];
        if (t0 == t1) continue;
        const PixPair P = pps[order[q]];
        const farm_raster_hdr &A = S.hdr[pa_ids[P.pair]];
        const farm_raster_hdr &B = S.hdr[pb_ids[P.pair]];
        TileDesc d;
        d.pair = P.pair;
        d.a_s0 = A.seq_start + S.pixoff[A.pixoff_start + P.pa];
        d.a_s1 = A.seq_start + S.pixoff[A.pixoff_start + P.pa + 1];
        d.b_s0 = B.seq_start + S.pixoff[B.pixoff_start + P.pb];
        d.b_s1 = B.seq_start + S.pixoff[B.pixoff_start + P.pb + 1];
        d.a_vert = A.vert_start;
        d.b_vert = B.vert_start;
        // the target pixel's sequences once: offsets, sequence entries, vertices
        const uint32_t nb = d.b_s1 - d.b_s0;
        d.rbytes = 16 + 8u * nb + 8u * (S.pixedges[B.pix_start + P.pb] + nb);
        d.pad[0] = d.pad[1] = d.pad[2] = 0;
        for (uint32_t t = t0; t < t1; t++) {
            d.tile = t - t0;
            map[t] = d;
        }
    }
}

// ------------------------------------------------------------------------
// Persistent refinement.  Each wave takes the next tile from a global
// cursor.  A tile is 64 lanes x W consecutive edge slots of one source
// pixel's sequences (consecutive lanes share one vertex), refined against
// the edge sequences of one target pixel, walked in chunks of C segments
// (C = 8, 16 or 32, chosen per call; the autotuner picks it).
// Per chunk every lane first tests its W edges against the chunk's box (an
// exact prune: no edge whose line keeps the whole box at a safe distance can
// meet a segment of the chunk under the reference predicate); the (lane,
// chunk) pairs that survive are queued in LDS, and every 64 queued items are
// evaluated one per lane: the lane's edges (read from LDS) against the
// chunk's segments (its own loads), with the screen below and, for the rare
// candidates, the exact reference predicate.
// ------------------------------------------------------------------------

// Boxes of the refinement's target chunks.  A sequence (first, len) is walked
// in chunks c = 0, 1, ..: steps base = 1 + C c .. min(len, base + C - 1), whose
// segments span vertices first + base - 1 .. first + base + n - 1.  The box is
// stored as a centre and half extents rounded up, so [c - h, c + h] holds
// every vertex of the chunk in exact arithmetic.  Infinite extents (the prune
// never applies) go to a box that reaches past 2^40, holds a non-finite
// vertex, or holds a segment of length 0: for c == d the reference's
// c.cross(a, d) is 0 for every a, so segment_intersect takes its collinear
// branch and answers with the bounding-box overlap alone, on either side of
// line ab.
__global__ __launch_bounds__(256) void k_chunk_count(DevSet S, int64_t npolys, uint32_t C,
                                                     uint32_t *__restrict__ cnt) {
    const int64_t p = blockIdx.x;
    if (p >= npolys) return;
    const farm_raster_hdr H = S.hdr[p];
    for (uint32_t s = threadIdx.x; s < H.nseqs; s += blockDim.x)
        cnt[H.seq_start + s] = (S.seqs[H.seq_start + s].y + C - 1) / C;
}

__global__ __launch_bounds__(256) void k_chunk_box(DevSet S, int64_t npolys, uint32_t C,
                                                   const uint32_t *__restrict__ choff, float4 *__restrict__ box) {
    const int64_t p = blockIdx.x;
    if (p >= npolys) return;
    const farm_raster_hdr H = S.hdr[p];
    const float *x = S.x + H.vert_start, *y = S.y + H.vert_start;
    for (uint32_t s = threadIdx.x; s < H.nseqs; s += blockDim.x) {
        const uint2 sq = S.seqs[H.seq_start + s];
        uint32_t o = choff[H.seq_start + s];
        for (uint32_t base = 1; base <= sq.y; base += C, o++) {
            const uint32_t n = min(C, sq.y - base + 1);
            float x0 = x[sq.x + base - 1], x1 = x0, y0 = y[sq.x + base - 1], y1 = y0;
            bool open = isfinite(x0) && isfinite(y0);  // false: never prune this chunk
            float lx = x0, ly = y0;
            for (uint32_t m = 0; m < n; m++) {
                const float vx = x[sq.x + base + m], vy = y[sq.x + base + m];
                open = open && isfinite(vx) && isfinite(vy) && !(vx == lx && vy == ly);
                lx = vx;
                ly = vy;
                x0 = fminf(x0, vx);
                x1 = fmaxf(x1, vx);
                y0 = fminf(y0, vy);
                y1 = fmaxf(y1, vy);
            }
            const float cx = 0.5f * x0 + 0.5f * x1, cy = 0.5f * y0 + 0.5f * y1;
            // a rounded difference is within 2^-24 of the exact one; the
            // factor and the smallest normal cover it and the product's rounding
            float hx = fmaxf(x1 - cx, cx - x0) * 1.0000003f + 1.1754944e-38f;
            float hy = fmaxf(y1 - cy, cy - y0) * 1.0000003f + 1.1754944e-38f;
            const float big = 0x1p40f;
            if (!(open && fabsf(cx) + hx < big && fabsf(cy) + hy < big)) hx = hy = INFINITY;
            box[o] = make_float4(cx, cy, hx, hy);
        }
    }
}

// Chunk boxes read through the constant address space: they do not change
// during a launch, and uniform addresses then become scalar loads.
typedef __attribute__((address_space(4))) const float cfloat;

constexpr uint32_t kRefQueue = 512;  // queued items per wave: a power of 2 above 63 + 64 kRefScreen

template <int W, int C>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_refine(DevSet S, const uint32_t *__restrict__ pa_all,
                                                 const uint32_t *__restrict__ pb_all,
                                                 const IsxDesc *__restrict__ desc,
                                                 const TileDesc *__restrict__ tile_map, uint64_t cap_pp,
                                                 uint64_t cap_tiles, uint32_t *__restrict__ cursor,
                                                 uint32_t *__restrict__ result, unsigned long long *__restrict__ ctr,
                                                 int prune) {
    // the tile total, unless a buffer of an earlier phase overflowed (the
    // host then grows the workspace and repeats the call, or fails it)
    if (ctr[kCtrPixPairs] > cap_pp || ctr[kCtrTiles] > cap_tiles) return;
    const uint32_t total = (uint32_t)ctr[kCtrTiles];
    uint32_t *pair_tests = desc->pair_tests;
    // per wave: the tile's source slots (x then y, W + 1 each) and edge bits
    // of every lane, and the ring of queued items (first vertex of the target
    // chunk, source lane | steps << 8)
    __shared__ float s_src[4][64][2 * (W + 1)];
    __shared__ uint32_t s_ev[4][64];
    __shared__ uint2 s_ring[4][kRefQueue];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    float(*src)[2 * (W + 1)] = s_src[wv];
    uint32_t *sev = s_ev[wv];
    uint2 *ring = s_ring[wv];
    // work counters: per tile in 32-bit lane registers, summed over the wave
    // at the end of the tile into this wave's 64-bit totals in LDS (64-bit
    // lane totals live across the whole loop and push the kernel past its
    // register budget)
    __shared__ unsigned long long s_acc[4][3];  // tests, refine bytes, pruned tests
    if (lane < 3) s_acc[wv][lane] = 0;
    unsigned long long items = 0;
    for (;;) {
        uint32_t tests = 0, rbytes = 0, pruned = 0;
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(cursor, 1u);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= total) break;
        // the tile's descriptor (one uniform load), then its pair's flag
        const TileDesc td = tile_map[t];
        const uint32_t pair = __builtin_amdgcn_readfirstlane(td.pair);
        const uint32_t tile = __builtin_amdgcn_readfirstlane(td.tile);
        if (__hip_atomic_load(&result[pair], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
        items++;
        const uint2 *sqa = S.seqs;  // absolute sequence indices
        const uint2 *sqb = S.seqs;
        const float *ax_ = S.x + __builtin_amdgcn_readfirstlane(td.a_vert);
        const float *ay_ = S.y + __builtin_amdgcn_readfirstlane(td.a_vert);
        const float4 *bq_ = S.xyd + __builtin_amdgcn_readfirstlane(td.b_vert);
        const uint32_t *bchoff = S.choff;
        const uint32_t a_s0 = __builtin_amdgcn_readfirstlane(td.a_s0), a_s1 = __builtin_amdgcn_readfirstlane(td.a_s1);
        const uint32_t b_s0 = __builtin_amdgcn_readfirstlane(td.b_s0);
        const uint32_t b_s1 = __builtin_amdgcn_readfirstlane(td.b_s1);
        // the tile's body, invoked once below (as a lambda it is allocated
        // to fewer registers than as a plain block: 128 against 155 VGPRs at
        // W = 8, four waves per SIMD instead of three)
        auto run = [&](auto width_tag) {
            constexpr int E = decltype(width_tag)::value;
            // source vertices of this lane: slots first .. first + E
            const uint32_t first = tile * 64u * W + (uint32_t)lane * E;
            float px[E + 1], py[E + 1];
            bool ev[E];
            const uint32_t nsq = a_s1 - a_s0;
            if (nsq <= 64) {
                // every lane loads one of the pixel's sequences and its first
                // slot (a wave prefix sum of len + 1) into the wave's LDS slot
                // array (free until this tile's slots are published); a slot's
                // sequence is then found by a binary search there and the
                // lane's later slots by stepping on (LDS reads only, so lanes
                // may diverge)
                const uint2 mine = (uint32_t)lane < nsq ? sqa[a_s0 + lane] : make_uint2(0u, 0u);
                uint32_t nslots;
                const uint32_t sstart = wave_excl_scan((uint32_t)lane < nsq ? mine.y + 1 : 0u, lane, &nslots);
                src[lane][0] = __uint_as_float(mine.x);  // (bit copies: the array holds floats)
                src[lane][1] = __uint_as_float(mine.y);
                src[lane][2] = __uint_as_float(sstart);
                wave_sync();
                auto ent = [&](int j, int f) { return __float_as_uint(src[j][f]); };
                int k = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int c = k + step;
                    if (c < (int)nsq && ent(c, 2) <= first) k = c;
                }
                uint32_t sbase = ent(k, 2);
                uint2 sq = make_uint2(ent(k, 0), ent(k, 1));
                for (int m = 0; m <= E; m++) {
                    const uint32_t s = first + m;
                    while (k + 1 < (int)nsq && s >= sbase + sq.y + 1) {
                        sbase += sq.y + 1;
                        k++;
                        sq = make_uint2(ent(k, 0), ent(k, 1));
                    }
                    uint32_t o = s - sbase;
                    const bool inside = s < sbase + sq.y + 1;
                    if (!inside) o = sq.y;  // past the last slot: clamp (no edge)
                    const uint32_t v = sq.x + o;
                    px[m] = ax_[v];
                    py[m] = ay_[v];
                    if (m < E) ev[m] = inside && o < sq.y;
                }
                wave_sync();  // every lane's reads done before the slots overwrite the array
            } else {
                uint32_t k = a_s0, sbase = 0;
                uint2 sq = sqa[k];
                for (int m = 0; m <= E; m++) {
                    const uint32_t s = first + m;
                    while (k + 1 < a_s1 && s >= sbase + sq.y + 1) {
                        sbase += sq.y + 1;
                        k++;
                        sq = sqa[k];
                    }
                    uint32_t o = s - sbase;
                    const bool inside = s < sbase + sq.y + 1;
                    if (!inside) o = sq.y;  // past the last slot: clamp (no edge)
                    const uint32_t v = sq.x + o;
                    px[m] = ax_[v];
                    py[m] = ay_[v];
                    if (m < E) ev[m] = inside && o < sq.y;
                }
            }
            int nvalid = 0;
            uint32_t evb = 0;
#pragma unroll
            for (int e = 0; e < E; e++) {
                nvalid += ev[e] ? 1 : 0;
                evb |= (ev[e] ? 1u : 0u) << e;
            }
            // publish this lane's slots for the lanes that evaluate its items
#pragma unroll
            for (int v = 0; v <= E; v++) {
                src[lane][v] = px[v];
                src[lane][E + 1 + v] = py[v];
            }
            sev[lane] = evb;
            __builtin_amdgcn_wave_barrier();
            bool hit = false;
            bool stop = false;
            uint32_t qhead = 0, qn = 0, screened = 0;  // wave-uniform
            uint32_t scr_steps = 0, queued_steps = 0;  // target steps screened (uniform) / queued by this lane

            // Evaluate the n <= 64 oldest queued items, one per lane: the source
            // lane's E edges against the chunk's nstep target segments.  Per
            // target vertex q and source edge (a, b):
            //   O(q) = a.cross(b, q)   -- o1/o2 of consecutive steps
            //   o(p) = c.cross(p, d)   -- per source vertex p, for the segment (c, d)
            // carried from step to step: tx/ty = p - c serves O(c) and, by the
            // antisymmetry of rounded subtraction, o(p) of the next step.  A
            // superset filter (o1*o2 <= 0 or o3 == 0, and o3*o4 <= 0) sends
            // the rare candidates to the exact reference predicate.
            //
            // Fewer than 64 items (a tile's last) are split over the lanes: k =
            // 64 / n rounded down to a power of two (at most C) lanes per item,
            // part j taking its steps j C / k .. (j + 1) C / k - 1.  A part is
            // the item's segments from its own first vertex on, so the split
            // evaluates the same tests in C / k steps instead of C.
            auto drain = [&](uint32_t n) {
                uint32_t lk = 0;  // log2 k (uniform)
                while ((1u << lk) < (uint32_t)C && (n << (lk + 1)) <= 64u) lk++;
                const uint32_t seg = (uint32_t)C >> lk;
                const uint32_t item = (uint32_t)lane >> lk, m0 = ((uint32_t)lane & ((1u << lk) - 1u)) * seg;
                const uint2 it = ring[(qhead + item) & (kRefQueue - 1)];
                const uint32_t nsf = item < n ? (it.y >> 8) : 0u;
                const uint32_t ns = nsf > m0 ? min(seg, nsf - m0) : 0u;
                const bool has = ns > 0u;
                const uint32_t sl = has ? (it.y & 63u) : (uint32_t)lane;
                const uint32_t vs = has ? it.x + m0 : 0u;  // the part's first vertex c
                float qpx[E + 1], qpy[E + 1];
#pragma unroll
                for (int v = 0; v <= E; v++) {
                    qpx[v] = src[sl][v];
                    qpy[v] = src[sl][E + 1 + v];
                }
                const uint32_t eb = sev[sl];
                float qbax[E], qbay[E], qinv[E];
                int qnv = 0;
#pragma unroll
                for (int e = 0; e < E; e++) {
                    qbax[e] = qpx[e + 1] - qpx[e];
                    qbay[e] = qpy[e + 1] - qpy[e];
                    const bool ve = (eb >> e) & 1u;
                    qinv[e] = ve ? -INFINITY : INFINITY;  // lifts a slot without an edge out of the screen
                    qnv += ve ? 1 : 0;
                }
                const float4 c04 = bq_[vs];
                const float c0x = c04.x, c0y = c04.y;
                float tx[E + 1], ty[E + 1], Op[E];
#pragma unroll
                for (int v = 0; v <= E; v++) {
                    tx[v] = qpx[v] - c0x;
                    ty[v] = qpy[v] - c0y;
                }
#pragma unroll
                for (int e = 0; e < E; e++) {
                    float l = qbay[e] * tx[e];
                    float r = qbax[e] * ty[e];
                    Op[e] = l - r;  // a.cross(b, c), bit-identical to the reference form
                }
                // running minimum of the screen, even and odd edges apart
                float gmin[2] = {INFINITY, INFINITY};
#pragma unroll 4
                for (int m = 0; m < C; m++) {
                    if ((uint32_t)m < ns) {
                        // the step's end q and direction (q - c), precomputed with
                        // the same rounding, in one load
                        const float4 q4 = bq_[vs + 1 + m];
                        const float qx = q4.x, qy = q4.y, ddx = q4.z, ddy = q4.w;
                        float o[E + 1];
#pragma unroll
                        for (int v = 0; v <= E; v++) {
                            float l = tx[v] * ddy;
                            float r = ty[v] * ddx;
                            o[v] = l - r;  // c.cross(p_v, d)
                            tx[v] = qpx[v] - qx;
                            ty[v] = qpy[v] - qy;
                        }
#pragma unroll
                        for (int e = 0; e < E; e++) {
                            float l = qbay[e] * tx[e];
                            float r = qbax[e] * ty[e];
                            const float On = l - r;  // a.cross(b, d)
                            // superset of segment_intersect: both pairs of
                            // orientations have opposite (or zero) signs, or the
                            // collinear branch may apply (o3 == 0).  A slot
                            // without an edge enters only through o3 == 0, a
                            // false candidate the exact replay discards.
                            const float h = fmaxf(fmaxf(o[e] * o[e + 1], Op[e] * On), qinv[e]);
                            gmin[e & 1] = fminf(fminf(gmin[e & 1], h), fabsf(o[e]));
                            Op[e] = On;
                        }
                    }
                }
                const bool maybe = fminf(gmin[0], gmin[1]) <= 0.f;
                if (__any(maybe)) {
                    // rare: re-evaluate the item with the reference predicate
                    if (maybe) {
                        float ccx = c0x, ccy = c0y;
                        for (uint32_t m = 0; m < ns; m++) {
                            const float4 d4 = bq_[vs + 1 + m];
                            const float dx = d4.x, dy = d4.y;
#pragma unroll
                            for (int e = 0; e < E; e++)
                                if (((eb >> e) & 1u) &&
                                    seg_intersect(qpx[e], qpy[e], qpx[e + 1], qpy[e + 1], ccx, ccy, dx, dy))
                                    hit = true;
                            ccx = dx;
                            ccy = dy;
                        }
                    }
                }
                tests += (uint32_t)qnv * ns;
                qhead = (qhead + n) & (kRefQueue - 1);
                qn -= n;
                if (__any(hit) || __hip_atomic_load(&result[pair], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    stop = true;
            };

            for (uint32_t j = b_s0; j < b_s1 && !stop; j++) {
                const uint2 sq = sqb[j];
                const uint32_t v0 = __builtin_amdgcn_readfirstlane(sq.x);
                const uint32_t L = __builtin_amdgcn_readfirstlane(sq.y);
                const cfloat *cbox = (const cfloat *)(S.chbox + __builtin_amdgcn_readfirstlane(bchoff[j]));
                const uint32_t nch = (L + C - 1) / C;
                for (uint32_t g = 0; g < nch && !stop; g += kRefScreen) {
                    // Exact prune of this lane's edges against kRefScreen chunks
                    // (their boxes are scalar loads; boxes past the sequence's
                    // last chunk are read, the array keeps spare entries, but not
                    // used).  q >= Q, the L1 reach from any vertex of the lane's
                    // edges to the chunk box, is taken from the box of those
                    // vertices.  If every box point p has
                    // |a.cross(b, p)| > M_lin = 2^-18 q |b - a|_1 in exact
                    // arithmetic, the reference's f32 a.cross(b, p) has one
                    // nonzero sign on the whole chunk (rounding <= 3 2^-24 Q
                    // |b - a|_1), so its sign test is false for every segment of
                    // the chunk.  Its collinear branch (c.cross(a, d) ==
                    // c.cross(b, d) == 0) answers with inter1's exact bounding-box
                    // overlap, impossible when the chunk box and the box of the
                    // lane's vertices are apart on an axis; otherwise it needs a
                    // and b within rounding of line cd, which puts c within
                    // 2^-20.7 Q (|b - a|_1 + 2Q) of line ab, and the margin is
                    // M = 2^-18 q (|b - a|_1 + 2q).  Both margins get 2^-80 added.
                    // The cross product at the box centre is evaluated relative
                    // to the lane's first vertex o: C = s.cross(a - o) per edge
                    // (s = b - a), then s.cross(c - o) - C in two fmas; with the
                    // spread over the box and M in two more fmas the f32
                    // evaluation errs by at most about 13 2^-24 q |b - a|_1
                    // (every operand is within 2q of o), which M covers twice
                    // over.  (A chunk with a segment of length 0 has an
                    // infinite box and is never pruned.)
                    // this lane's slots again, from LDS (not kept in registers
                    // across the evaluation of queued items)
                    float sbx[E], sby[E], scc[E];
                    float lxl, lxh, lyl, lyh;  // box of the lane's slot vertices
                    const float ox = src[lane][0], oy = src[lane][E + 1];
                    {
                        float nx = ox, ny = oy;
                        lxl = lxh = nx;
                        lyl = lyh = ny;
#pragma unroll
                        for (int e = 0; e < E; e++) {
                            const float ax = nx, ay = ny;
                            nx = src[lane][e + 1];
                            ny = src[lane][E + 2 + e];
                            lxl = fminf(lxl, nx);
                            lxh = fmaxf(lxh, nx);
                            lyl = fminf(lyl, ny);
                            lyh = fmaxf(lyh, ny);
                            const bool ve = (evb >> e) & 1u;
                            // a slot without an edge gets s = 0 and C = +inf: its
                            // centre cross is -inf and the box is "pruned" for it
                            // (an infinite box makes the threshold NaN instead,
                            // which queues the item; the evaluation skips the slot)
                            sbx[e] = ve ? nx - ax : 0.f;
                            sby[e] = ve ? ny - ay : 0.f;
                            const float axo = ax - ox, ayo = ay - oy;
                            scc[e] = ve ? fmaf(sbx[e], ayo, -(sby[e] * axo)) : INFINITY;
                        }
                    }
                    bool need[kRefScreen];
#pragma unroll
                    for (int i = 0; i < (int)kRefScreen; i++) {
                        const float4 bb = make_float4(cbox[4 * (g + i)], cbox[4 * (g + i) + 1],
                                                      cbox[4 * (g + i) + 2], cbox[4 * (g + i) + 3]);
                        // the reach q (rounded up), and whether the boxes are apart
                        // on an axis (by more than the f32 rounding of the tests)
                        const float ex0 = bb.x - lxl, ex1 = bb.x - lxh;
                        const float ey0 = bb.y - lyl, ey1 = bb.y - lyh;
                        const float q = ((fmaxf(fabsf(ex0), fabsf(ex1)) + fmaxf(fabsf(ey0), fabsf(ey1))) +
                                         (bb.z + bb.w)) * 1.000001f;
                        const float hx1 = bb.z * 1.000001f, hy1 = bb.w * 1.000001f;
                        // (ex1 > hx1 or -ex0 > hx1 or ey1 > hy1 or -ey0 > hy1, as two
                        // maxima and two compares)
                        const bool apart = (fmaxf(ex1, -ex0) > hx1) | (fmaxf(ey1, -ey0) > hy1);
                        // (without the prune: the largest finite margin keeps every
                        // chunk for every slot with an edge)
                        const float k1 = !prune ? 0x1.fffffep127f : q * 0x1p-18f;
                        const float k2 = !prune ? 0x1.fffffep127f : (apart ? 0x1p-80f : fmaf(q, q * 0x1p-17f, 0x1p-80f));
                        // |s_y| hx + |s_x| hy + (|s_x| + |s_y|) k1 + k2 as
                        // |s_y| (hx + k1) + |s_x| (hy + k1) + k2
                        const float hxk = bb.z + k1, hyk = bb.w + k1;
                        const float cxo = bb.x - ox, cyo = bb.y - oy;
                        need[i] = false;
#pragma unroll
                        for (int e = 0; e < E; e++) {
                            const float oc = fmaf(sbx[e], cyo, -fmaf(sby[e], cxo, scc[e]));  // at the box centre
                            const float thr = fmaf(fabsf(sby[e]), hxk, fmaf(fabsf(sbx[e]), hyk, k2));
                            need[i] = need[i] | !(fabsf(oc) > thr);  // bitwise: no branch per edge
                        }
                    }
#pragma unroll
                    for (int i = 0; i < (int)kRefScreen; i++) {
                        const uint32_t c = g + i;
                        if (c >= nch) break;
                        const uint32_t nstep = min((uint32_t)C, L - c * C);
                        scr_steps += nstep;
                        const unsigned long long bm = __ballot(need[i]);
                        if (bm) {
                            // append this chunk for every lane that needs it
                            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                            if (need[i]) {
                                ring[(qhead + qn + below) & (kRefQueue - 1)] =
                                    make_uint2(v0 + c * C, (uint32_t)lane | (nstep << 8));
                                queued_steps += nstep;
                            }
                            qn += (uint32_t)__popcll(bm);
                        }
                    }
                    while (qn >= 64 && !stop) {
                        __builtin_amdgcn_wave_barrier();
                        drain(64);
                    }
                    if (!stop && (++screened & 7u) == 0 &&
                        __hip_atomic_load(&result[pair], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        stop = true;
                }
            }
            if (!stop && qn) {
                __builtin_amdgcn_wave_barrier();
                drain(qn);
            }
            if (__any(hit)) {
                if (lane == 0) __hip_atomic_fetch_or(&result[pair], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }

            // algorithmic refine bytes of this tile: its source vertices plus
            // the target pixel's sequences once
            if (lane == 0) rbytes += __builtin_amdgcn_readfirstlane(td.rbytes);
            rbytes += (uint32_t)nvalid * 8u;
            // tests of this lane's edges the prune skipped: every screened step
            // that was not queued
            pruned += (uint32_t)nvalid * (scr_steps - queued_steps);
        };
        run(std::integral_constant<int, W>());
        // the tile's totals (a tile holds at most 512 edges against one target
        // pixel's: its sums stay far below 2^64 and each lane's below 2^32)
        unsigned long long t64 = tests, b64 = rbytes, p64 = pruned;
        for (int o = 32; o > 0; o >>= 1) {
            t64 += __shfl_xor(t64, o);
            b64 += __shfl_xor(b64, o);
            p64 += __shfl_xor(p64, o);
        }
        if (lane == 0) {
            s_acc[wv][0] += t64;
            s_acc[wv][1] += b64;
            s_acc[wv][2] += p64;
            // executed tests of this tile, for the cost labels
            if (pair_tests) atomicAdd(&pair_tests[pair], (uint32_t)min(t64, 0xffffffffull));
        }
    }
    if (lane == 0) {
        atomicAdd(&ctr[kCtrTests], s_acc[wv][0]);
        atomicAdd(&ctr[kCtrItems], items);
        atomicAdd(&ctr[kCtrRefineBytes], s_acc[wv][1]);
        atomicAdd(&ctr[kCtrPruned], s_acc[wv][2]);
    }
}

// ------------------------------------------------------------------------
// Plain kernel: one workgroup (256 lanes) per candidate pair, the direct
// formulation with no pruning and no queue.  The window is scanned for IN
// overlaps by all lanes; then each BORDER x BORDER pixel pair is refined in
// full: both pixels' edges are flattened into LDS (a block scan over the
// sequence lengths, then every lane fills edges, finding its sequence by a
// binary search over the scanned starts), as coordinates, up to kPlainCap
// edges per side at a time, and the lanes sweep the source x target edge
// grid in 16 x 16 blocks (lane t tests source edge t / 16 and target edge
// t % 16 of each block), with a block-wide early exit every 8 blocks.
// ------------------------------------------------------------------------
constexpr int kPlainCap = 1024;   // edges per side held in LDS per chunk
constexpr int kPlainSeqs = 256;   // sequences per pixel scanned per pass

// Fills dst[0 .. len) with the coordinates (a.x, a.y, b.x, b.y) of edges
// first .. first + len of the concatenation of a pixel's sequences
// sq[0 .. nseq), seq_off[k] the scanned start of sequence k.
__device__ void plain_fill(const uint2 *__restrict__ sq, const uint32_t *seq_off, uint32_t nseq, uint32_t first,
                           uint32_t len, const float *__restrict__ X, const float *__restrict__ Y, float4 *dst) {
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
        const uint32_t g = first + i;
        uint32_t lo = 0, hi = nseq;  // last k with seq_off[k] <= g
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (seq_off[mid] <= g) lo = mid;
            else hi = mid;
        }
        const uint32_t v = sq[lo].x + (g - seq_off[lo]);
        dst[i] = make_float4(X[v], Y[v], X[v + 1], Y[v + 1]);
    }
}

// Exclusive block scan of the lengths of sequences sq[0 .. n) (n <=
// kPlainSeqs): seq_off[k] for k < n, returns the total.
__device__ uint32_t plain_scan(const uint2 *__restrict__ sq, uint32_t n, uint32_t *seq_off, uint32_t *wsum) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t v = (uint32_t)t < n ? sq[t].y : 0u;
    uint32_t tot;
    const uint32_t ex = wave_excl_scan(v, lane, &tot);
    if (lane == 0) wsum[w] = tot;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int i = 0; i < 4; i++) {
        before += i < w ? wsum[i] : 0u;
        all += wsum[i];
    }
    if ((uint32_t)t < n) seq_off[t] = before + ex;
    __syncthreads();
    return all;
}

__global__ __launch_bounds__(256) void k_plain(DevSet S, const uint32_t *__restrict__ pa_ids,
                                                const uint32_t *__restrict__ pb_ids, uint32_t npairs,
                                                uint32_t *__restrict__ result, unsigned long long *__restrict__ ctr) {
    __shared__ float4 ea[kPlainCap];
    __shared__ float4 eb[kPlainCap];
    __shared__ uint32_t offa[kPlainSeqs], offb[kPlainSeqs], wsum[4];
    __shared__ int s_flag;
    const uint32_t k = blockIdx.x;
    if (k >= npairs) return;
    const int tid = threadIdx.x;
    const farm_raster_hdr A = S.hdr[pa_ids[k]];
    const farm_raster_hdr B = S.hdr[pb_ids[k]];
    const uint8_t *sa = S.status + A.pix_start;
    const uint8_t *sb = S.status + B.pix_start;
    const uint32_t *pofa = S.pixoff + A.pixoff_start;
    const uint32_t *pofb = S.pixoff + B.pixoff_start;
    const uint2 *sqa = S.seqs + A.seq_start;
    const uint2 *sqb = S.seqs + B.seq_start;
    const float *AX = S.x + A.vert_start, *AY = S.y + A.vert_start;
    const float *BX = S.x + B.vert_start, *BY = S.y + B.vert_start;
    const Window wa = window_of(A, B.mbr[0], B.mbr[1], B.mbr[2], B.mbr[3]);
    const int nwx = wa.x1 - wa.x0 + 1, nwy = wa.y1 - wa.y0 + 1;
    const int nwa = nwx * nwy;
    if (tid == 0) s_flag = 0;
    __syncthreads();
    // IN overlap scan: lanes over source window pixels
    for (int w = tid; w < nwa; w += blockDim.x) {
        const int x = wa.x0 + w / nwy, y = wa.y0 + w % nwy;
        const int st_a = show(sa[y * A.dimx + x]);
        if (st_a == 0) continue;
        double lx, ly, hx, hy;
        pixel_box(A, x, y, lx, ly, hx, hy);
        const Window wb = window_of(B, lx, ly, hx, hy);
        for (int u = wb.x0; u <= wb.x1; u++)
            for (int v = wb.y0; v <= wb.y1; v++) {
                const int st_b = show(sb[v * B.dimx + u]);
                if (st_b != 0 && (st_a == 2 || st_b == 2)) s_flag = 1;
            }
    }
    __syncthreads();
    if (s_flag) {
        if (tid == 0) {
            result[k] = 1;
            atomicAdd(&ctr[kCtrInHits], 1ull);
        }
        return;
    }
    unsigned long long tests = 0;
    bool hit = false;
    const int ti = tid >> 4, tj = tid & 15;  // this lane's place in a 16 x 16 block
    // refinement: uniform loop over the BORDER x BORDER pixel pairs
    for (int w = 0; w < nwa && !hit; w++) {
        const int x = wa.x0 + w / nwy, y = wa.y0 + w % nwy;
        const int pa = y * A.dimx + x;
        if (show(sa[pa]) != 1) continue;
        double lx, ly, hx, hy;
        pixel_box(A, x, y, lx, ly, hx, hy);
        const Window wb = window_of(B, lx, ly, hx, hy);
        for (int u = wb.x0; u <= wb.x1 && !hit; u++)
            for (int v = wb.y0; v <= wb.y1 && !hit; v++) {
                const int pb = v * B.dimx + u;
                if (show(sb[pb]) != 1) continue;
                // the pixels' sequences, kPlainSeqs at a time on each side
                for (uint32_t s0a = pofa[pa]; s0a < pofa[pa + 1] && !hit; s0a += kPlainSeqs) {
                    const uint32_t nsa = min((uint32_t)kPlainSeqs, pofa[pa + 1] - s0a);
                    const uint32_t na = plain_scan(sqa + s0a, nsa, offa, wsum);
                    for (uint32_t s0b = pofb[pb]; s0b < pofb[pb + 1] && !hit; s0b += kPlainSeqs) {
                        const uint32_t nsb = min((uint32_t)kPlainSeqs, pofb[pb + 1] - s0b);
                        const uint32_t nb = plain_scan(sqb + s0b, nsb, offb, wsum);
                        for (uint32_t ca = 0; ca < na && !hit; ca += kPlainCap) {
                            const uint32_t la = min((uint32_t)kPlainCap, na - ca);
                            __syncthreads();
                            plain_fill(sqa + s0a, offa, nsa, ca, la, AX, AY, ea);
                            for (uint32_t cb = 0; cb < nb && !hit; cb += kPlainCap) {
                                const uint32_t lb = min((uint32_t)kPlainCap, nb - cb);
                                __syncthreads();
                                plain_fill(sqb + s0b, offb, nsb, cb, lb, BX, BY, eb);
                                __syncthreads();
                                const uint32_t bi = (la + 15) / 16, bj = (lb + 15) / 16;
                                uint32_t blk = 0;
                                bool done = false;  // block-uniform: set only from __syncthreads_or
                                for (uint32_t i0 = 0; i0 < bi && !done; i0++) {
                                    const uint32_t i = 16 * i0 + ti;
                                    const float4 a = ea[min(i, la - 1)];
                                    for (uint32_t j0 = 0; j0 < bj; j0++, blk++) {
                                        const uint32_t j = 16 * j0 + tj;
                                        if (i < la && j < lb) {
                                            const float4 b = eb[j];
                                            tests++;
                                            if (seg_intersect(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w)) hit = true;
                                        }
                                        if ((blk & 7) == 7 && __syncthreads_or(hit)) {
                                            done = true;
                                            break;
                                        }
                                    }
                                }
                                hit = __syncthreads_or(hit);
                            }
                        }
                    }
                }
            }
    }
    for (int o = 32; o > 0; o >>= 1) tests += __shfl_xor(tests, o);
    if ((tid & 63) == 0) atomicAdd(&ctr[kCtrTests], tests);
    if (tid == 0 && hit) result[k] = 1;
}

// xyd[v] = (x[v], y[v], x[v] - x[v - 1], y[v] - y[v - 1]) over n padded
// vertices: the vertex and the same rounded f32 difference the refinement
// would form for the step arriving at it
__global__ __launch_bounds__(256) void k_vdiff(const float *__restrict__ x, const float *__restrict__ y, uint64_t n,
                                               float4 *__restrict__ xyd) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float vx = x[v], vy = y[v];
    xyd[v] = v == 0 ? make_float4(vx, vy, 0.f, 0.f) : make_float4(vx, vy, vx - x[v - 1], vy - y[v - 1]);
}

// per-pixel edge count of every BORDER pixel (sum of its sequence lengths)
__global__ __launch_bounds__(256) void k_pixedges(DevSet S, int64_t npolys, uint32_t *__restrict__ out) {
    const int64_t p = blockIdx.x;
    if (p >= npolys) return;
    const farm_raster_hdr H = S.hdr[p];
    const int n = H.dimx * H.dimy;
    const uint32_t *pof = S.pixoff + H.pixoff_start;
    const uint2 *sq = S.seqs + H.seq_start;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t c = 0;
        if (show(S.status[H.pix_start + i]) == 1)
            for (uint32_t s = pof[i]; s < pof[i + 1]; s++) c += sq[s].y;
        out[H.pix_start + i] = c;
    }
}

__global__ __launch_bounds__(256) void k_finalize(const uint32_t *__restrict__ result,
                                                   const IsxDesc *__restrict__ desc,
                                                   unsigned long long *__restrict__ ctr) {
    const uint64_t n = desc->n;
    uint8_t *out = desc->out;
    unsigned long long f = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = result[i] ? 1u : 0u;
        if (out) out[i] = (uint8_t)r;
        f += r;
    }
    for (int o = 32; o > 0; o >>= 1) f += __shfl_xor(f, o);
    if ((threadIdx.x & 63) == 0 && f) atomicAdd(&ctr[kCtrFound], f);
}

#define FARM_TILES_INST(EE)                                                                                    \
    template __global__ void k_tiles<EE>(const PixPair *, const uint32_t *, unsigned long long *, uint64_t,      \
                                         uint32_t *);
#define FARM_REFINE_INST(EE, CC)                                                                               \
    template __global__ void k_refine<EE, CC>(DevSet, const uint32_t *, const uint32_t *, const IsxDesc *,      \
                                              const TileDesc *, uint64_t, uint64_t, uint32_t *, uint32_t *,     \
                                              unsigned long long *, int);
FARM_TILES_INST(1)
FARM_TILES_INST(2)
FARM_TILES_INST(4)
FARM_TILES_INST(8)
FARM_REFINE_INST(1, 8)
FARM_REFINE_INST(2, 8)
FARM_REFINE_INST(4, 8)
FARM_REFINE_INST(8, 8)
FARM_REFINE_INST(1, 16)
FARM_REFINE_INST(2, 16)
FARM_REFINE_INST(4, 16)
FARM_REFINE_INST(8, 16)
FARM_REFINE_INST(1, 32)
FARM_REFINE_INST(2, 32)
FARM_REFINE_INST(4, 32)
FARM_REFINE_INST(8, 32)
#undef FARM_TILES_INST
#undef FARM_REFINE_INST

}  // namespace farm
