// Device scratch of one batched intersect call and the counters its kernels
// share.  Every region is carved from one block of device memory: a
// handle-owned allocation grown on demand, or the library's fixed scratch
// array for farm_intersect_noalloc.  The kernels read the pair window and the
// element counts of later phases from device memory (IsxDesc, ctr), so one
// captured launch sequence serves every window and the host never waits in
// the middle of a call.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_layout.h"

namespace farm {

// The pair window of a call, written by k_set_desc before the kernels run.
struct IsxDesc {
    unsigned long long begin;  // first pair of the window in the uploaded pair list
    unsigned long long n;      // pairs in the window
    uint8_t *out;              // 0/1 per pair, may be null
    uint32_t *pair_tests;      // executed segment tests per pair, may be null
};

// Counter slots (unsigned long long ctr[kCtrSlots]).
enum : int {
    kCtrFilterBytes = 0,  // algorithmic filter bytes
    kCtrInHits = 1,       // pairs decided by an IN overlap
    kCtrTests = 2,        // segment tests executed
    kCtrItems = 3,        // refinement tiles executed
    kCtrRefineBytes = 4,  // algorithmic refinement bytes
    kCtrFound = 5,        // pairs reported true
    kCtrTiles = 7,        // refinement tiles of the window (64-bit total)
    kCtrPruned = 8,       // segment tests skipped by the exact chunk prune
    kCtrPixPairs = 9,     // BORDER x BORDER pixel pairs emitted (64-bit total)
    kCtrSlots = 16,
};

// Scan partials: one per block of the three-kernel scan.
constexpr int kScanParts = 1024;

// What the refinement needs of one tile, gathered by k_tile_map so that the
// persistent kernel starts a tile with one load instead of a chain of
// dependent ones (order -> pixel pair -> pair ids -> headers -> pixel
// offsets).  Sequence indices are absolute (into DevSet::seqs / choff).
struct TileDesc {
    uint32_t pair;     // index in the batch
    uint32_t tile;     // tile number within its pixel pair
    uint32_t a_s0, a_s1;  // source pixel's edge sequences
    uint32_t b_s0, b_s1;  // target pixel's edge sequences
    uint32_t a_vert, b_vert;  // first vertex of the source / target polygon
    uint32_t rbytes;   // algorithmic refine bytes of the tile's target side
    uint32_t pad[3];
};
static_assert(sizeof(TileDesc) == 48, "tile descriptor layout");

struct IsxWork {
    IsxDesc *desc;
    unsigned long long *ctr;  // kCtrSlots
    uint32_t *hist;           // 64 cost buckets
    uint32_t *cursor;         // 64: bucket cursors, then the tile cursor
    uint32_t *part;           // kScanParts + 1
    uint32_t *result;         // cap_pairs
    uint32_t *npp;            // cap_pairs: BORDER x BORDER pixel pairs per pair
    uint32_t *pp_off;         // cap_pairs + 1
    uint8_t *bits;            // cap_pairs: staging of the 0/1 results for a host copy
    PixPair *pps;             // cap_pp
    uint8_t *bucket;          // cap_pp
    uint32_t *order;          // cap_pp
    uint32_t *tiles;          // cap_pp
    uint32_t *tile_off;       // cap_pp + 1
    TileDesc *tile_map;       // cap_tiles
    uint64_t cap_pairs = 0, cap_pp = 0, cap_tiles = 0;
};

constexpr uint64_t isx_align(uint64_t b) { return (b + 255) & ~255ull; }

// Bytes of a workspace with these capacities (every region 256-byte aligned).
constexpr uint64_t isx_bytes(uint64_t cap_pairs, uint64_t cap_pp, uint64_t cap_tiles) {
    return isx_align(sizeof(IsxDesc)) + isx_align(8ull * kCtrSlots) + isx_align(4ull * 64) + isx_align(4ull * 64) +
           isx_align(4ull * (kScanParts + 1)) + isx_align(4ull * cap_pairs) + isx_align(4ull * cap_pairs) +
           isx_align(4ull * (cap_pairs + 1)) + isx_align(cap_pairs) + isx_align(sizeof(PixPair) * cap_pp) +
           isx_align(cap_pp) + isx_align(4ull * cap_pp) + isx_align(4ull * cap_pp) + isx_align(4ull * (cap_pp + 1)) +
           isx_align(sizeof(TileDesc) * cap_tiles);
}

inline IsxWork isx_carve(void *base, uint64_t cap_pairs, uint64_t cap_pp, uint64_t cap_tiles) {
    IsxWork w;
    char *p = (char *)base;
    auto take = [&](uint64_t bytes) {
        char *r = p;
        p += isx_align(bytes);
        return (void *)r;
    };
    w.desc = (IsxDesc *)take(sizeof(IsxDesc));
    w.ctr = (unsigned long long *)take(8ull * kCtrSlots);
    w.hist = (uint32_t *)take(4ull * 64);
    w.cursor = (uint32_t *)take(4ull * 64);
    w.part = (uint32_t *)take(4ull * (kScanParts + 1));
    w.result = (uint32_t *)take(4ull * cap_pairs);
    w.npp = (uint32_t *)take(4ull * cap_pairs);
    w.pp_off = (uint32_t *)take(4ull * (cap_pairs + 1));
    w.bits = (uint8_t *)take(cap_pairs);
    w.pps = (PixPair *)take(sizeof(PixPair) * cap_pp);
    w.bucket = (uint8_t *)take(cap_pp);
    w.order = (uint32_t *)take(4ull * cap_pp);
    w.tiles = (uint32_t *)take(4ull * cap_pp);
    w.tile_off = (uint32_t *)take(4ull * (cap_pp + 1));
    w.tile_map = (TileDesc *)take(sizeof(TileDesc) * cap_tiles);
    w.cap_pairs = cap_pairs;
    w.cap_pp = cap_pp;
    w.cap_tiles = cap_tiles;
    return w;
}

// Element count of a phase, read on the device: the 64-bit total, or 0 when
// it exceeds the capacity of the buffers the phase writes (the host then
// grows the workspace, or reports FARM_ECAPACITY, and no kernel has written
// past a buffer).
__device__ __forceinline__ uint64_t bounded_count(const unsigned long long *src, uint64_t cap) {
    const unsigned long long v = *src;
    return v > cap ? 0ull : v;
}

}  // namespace farm
