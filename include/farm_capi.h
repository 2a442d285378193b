/*
 * farm_capi.h -- C ABI of the MI355X polygon-join library (libfarm.so).
 *
 * This is the drop-in boundary that replaces the reference's GPU entry points
 * declared in src/include/Ideal.h:205-213 (ResetDevice, cuda_create_buffer,
 * preprocess_for_gpu, cuda_intersect, cuda_intersection, cuda_within_polygon)
 * and the undeclared cuda_intersection_a (src/cuda/query_intersection.cu:756).
 * The reference passes a query_context* holding every host/device buffer
 * (src/include/query_context.h:171-229); here the same data crosses as plain
 * pointers and sizes, owned by an opaque handle.
 *
 * Conventions
 *  - Every entry point returns an int status: FARM_OK (0) or an FARM_E* code;
 *    farm_last_error() gives a message.  Nothing calls exit().
 *  - One HIP stream per handle; calls on one handle are serialised by the
 *    caller (not re-entrant).  Separate handles are independent.
 *  - Input arrays are host memory unless the name says "dev"; the library
 *    copies them and the caller keeps ownership.
 *  - Polygon ids: the uploaded set is one array; the reference GPU path also
 *    stores sources then targets in one array and shifts target ids by the
 *    number of sources (src/cuda/preprocessing.cu:32).
 */
#ifndef FARM_CAPI_H
#define FARM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FARM_OK 0
#define FARM_EINVAL 1     /* bad argument / inconsistent layout */
#define FARM_ENOMEM 2     /* device allocation failed */
#define FARM_EDEVICE 3    /* HIP runtime / launch error */
#define FARM_ECAPACITY 4  /* a bounded buffer would overflow */
#define FARM_ESTATE 5     /* call order violated (e.g. no polygons uploaded) */

/* Category count of the raster fill classes used by the CPU reference
 * (src/include/MyRaster.h:33): 0 = OUT, 15 = IN, 1..14 = BORDER. */
#define FARM_CATEGORY_COUNT 16

/*
 * Per-polygon raster header (the IDEAL model of one polygon).  Replaces the
 * reference's IdealOffset (src/include/Ideal.h:41-50) + RasterInfo
 * (src/include/MyRaster.h:14-21) pair.  128 bytes; the first 64 bytes are
 * what the raster filter reads.
 */
typedef struct farm_raster_hdr {
    double mbr[4];        /* raster box: low x, low y, high x, high y */
    double step[2];       /* pixel width / height */
    int32_t dimx, dimy;   /* pixel grid; pixel id = y * dimx + x */
    uint64_t pix_start;   /* first status byte in farm_polyset.status */
    /* --- refinement --- */
    uint64_t pixoff_start;  /* first entry in pixoff (dimx*dimy + 1 entries) */
    uint64_t seq_start;     /* first entry in seqs */
    uint64_t vert_start;    /* first vertex in x / y */
    uint32_t nverts;        /* ring length incl. the repeated first vertex */
    uint32_t nseqs;
    uint64_t vline_start;   /* first entry in voff (dimx + 2 entries) */
    uint64_t vnode_start;   /* first entry in vnodes */
    /* --- hierarchy (-h), replaces Ideal::layers (src/include/Ideal.h:66) --- */
    uint64_t layer_start;   /* first coarse layer in farm_polyset.layers */
    int32_t nlayers;        /* layers incl. this finest grid (coarse ones first); 0 = none */
    int32_t flags;          /* FARM_HDR_REJECTED: the reference raster builder aborts here */
} farm_raster_hdr;

#define FARM_HDR_REJECTED 1

/*
 * One coarse layer of a polygon's hierarchy (Hraster, src/include/Hraster.h):
 * a square-pixel grid and its fill classes.  Layer k (0 = coarsest) of
 * polygon i is layers[hdr[i].layer_start + k] for k < nlayers - 1; the
 * finest layer is the polygon's own raster.
 */
typedef struct farm_layer {
    double low[2];          /* grid origin (low x, low y) */
    double step;            /* pixel side */
    int32_t dimx, dimy;
    uint64_t status_start;  /* first byte in farm_polyset.lstatus */
} farm_layer;

/*
 * A rasterised polygon set in device layout (SoA of CSR), built on the host.
 *  status[pix_start + id]            fill class byte per pixel
 *  pixoff[pixoff_start + id .. +1]   edge-sequence range of pixel id, relative
 *                                    to seq_start (valid for BORDER pixels)
 *  seqs[2*(seq_start + k)]           first edge (relative to vert_start)
 *  seqs[2*(seq_start + k) + 1]       number of edges
 *  x/y[vert_start + v]               ring vertices (f32)
 *  voff/vnodes                       vertical grid-line crossings, per line
 *                                    sorted (used by point-in-polygon)
 */
typedef struct farm_polyset {
    int64_t npolys;
    const farm_raster_hdr *hdr;
    const uint8_t *status;
    uint64_t nstatus;
    const uint32_t *pixoff;
    uint64_t npixoff;
    const uint32_t *seqs; /* 2 * nseqs */
    uint64_t nseqs;
    const float *x;
    const float *y;
    uint64_t nverts;
    const uint32_t *voff;
    uint64_t nvoff;
    const double *vnodes;
    uint64_t nvnodes;
    /* hierarchy and per-pixel areas: needed by farm_within only (NULL / 0
     * for intersect-only sets) */
    const farm_layer *layers;
    uint64_t nlayers;
    const uint8_t *lstatus;
    uint64_t nlstatus;
    const double *areas;    /* clipped polygon area per finest pixel, like status */
    /* fill classes at category_count categories (-w b: 2^b, 4..256) for the
     * approximate mode and the learned features, indexed like status; NULL
     * (or category_count 0 / 16): the status bytes, 16 categories.  Exact
     * results never depend on it: OUT, BORDER and IN do not change. */
    const uint8_t *fill;
    int32_t category_count;
    int32_t reserved;
} farm_polyset;

/* Query modes of farm_intersect. */
#define FARM_MODE_EXACT_PLAIN 0    /* one candidate pair per workgroup */
#define FARM_MODE_EXACT_BATCHED 1  /* filter pass + cost-ordered persistent refine */
#define FARM_MODE_APPROX_HEURISTIC 2 /* reference -a: IN overlap or BORDER pair prob > 1 */
#define FARM_MODE_APPROX_LEARNED 3   /* MLP logit on pair features, threshold in opts */

typedef struct farm_opts {
    int32_t mode;
    int32_t tile_a;        /* source edges per lane of the batched refinement; 0 = autotuned */
    int32_t tile_b;        /* target segments per chunk box of the batched refinement (8, 16
                              or 32); 0 = autotuned.  A size other than the one the set's
                              boxes are built for rebuilds them (allocates, synchronises) */
    int32_t order;         /* batched refine queue order: 0 exact cost bound, 1 learned cost head */
    float approx_threshold;/* learned -a: logit above which a pair is accepted */
    float approx_reject;   /* learned -a: logit below which a pair is rejected; pairs in
                              [reject, threshold] fall back to exact refinement */
    float merge_threshold; /* -m, accepted for compatibility; exact results do not depend on it */
    int32_t flags;         /* FARM_FLAG_*; 0 for the defaults */
} farm_opts;

/* farm_opts.flags: refine every (source edge, target chunk) of the batched
 * kernel in full, without the exact chunk-box prune (the same results; for
 * measuring what the prune saves). */
#define FARM_FLAG_NO_PRUNE 1

/* Counters of the work actually executed (device counters). */
typedef struct farm_stats {
    uint64_t pairs;            /* candidate pairs processed */
    uint64_t in_hits;          /* pairs decided by an IN pixel overlap */
    uint64_t bb_pixel_pairs;   /* BORDER x BORDER pixel pairs enumerated */
    uint64_t refine_items;     /* refinement work items executed */
    uint64_t segment_tests;    /* segment x segment tests executed */
    uint64_t filter_bytes;     /* status and header bytes the filter read (a decided pair stops early) */
    uint64_t refine_bytes;     /* algorithmic refinement bytes */
    uint64_t found;            /* pairs reported true */
    double filter_ms;          /* device time of the filter kernel(s) */
    double refine_ms;          /* device time of the refinement kernel(s) */
    double total_ms;           /* device time of the whole call */
    uint64_t pruned_tests;     /* segment tests skipped whole by the exact chunk-box prune */
} farm_stats;

typedef struct farm_handle_s *farm_handle;

/* Open a handle on HIP device `device` (replaces ResetDevice +
 * cuda_create_buffer, src/cuda/preprocessing.cu:5-126,362-379). */
int farm_open(int device, farm_handle *out);
int farm_close(farm_handle h);
const char *farm_last_error(farm_handle h);

/* Upload a rasterised polygon set (replaces preprocess_for_gpu's packing and
 * host-to-device copies, src/cuda/preprocessing.cu:128-343). */
int farm_upload_polygons(farm_handle h, const farm_polyset *ps);

/* Rings of a polygon set before rasterization: ring i is x/y[vert_off[i] ..
 * vert_off[i+1]), its first vertex repeated at the end (the decoded ring of
 * VertexSequence, vertex_sequence.cpp:206-218), 2 <= length < 2^31. */
typedef struct farm_ringset {
    int64_t npolys;
    const uint64_t *vert_off; /* npolys + 1 */
    const float *x;
    const float *y;
} farm_ringset;

/* Array lengths of the set farm_rasterize built (farm_polyset's counts) and
 * the polygons the reference raster builder rejects. */
typedef struct farm_raster_info {
    int64_t nrejected;
    int64_t first_rejected; /* -1 if none */
    uint64_t nverts, nstatus, npixoff, nseqs, nvoff, nvnodes, nlayers, nlstatus;
    double build_ms;        /* device time from the uploaded vertices to the finished set */
    double upload_ms;       /* host wall time of the vertex upload */
} farm_raster_info;

/* Rasterize every ring on the device, vpr vertices per pixel (-v), and make
 * the result the handle's polygon set: the bytes farm_hostset_rasterize
 * (nlow = 0) or farm_hostset_rasterize_h (nlow > 0: -h with --NLow nlow)
 * gives, without the host raster build or upload.  Replaces preprocess /
 * process_rasterization (src/geometry/preprocessing.cpp:4-87,
 * Ideal::rasterization src/geometry/Ideal.cpp:920-955, with
 * evaluate_edges, process_crosses, process_intersection,
 * scanline_reandering and calculate_fullness, Ideal.cpp:35-918; -h:
 * MyRaster::grid_align MyRaster.cpp:444-491, Ideal::layering and
 * merge_status Ideal.cpp:1008-1099) and preprocess_for_gpu
 * (src/cuda/preprocessing.cu:128-343).  A rejected polygon (where the
 * reference builder asserts or never ends) becomes an empty raster flagged
 * FARM_HDR_REJECTED.  category_count is 16. */
int farm_rasterize(farm_handle h, const farm_ringset *rs, int vpr, int nlow, farm_raster_info *info);

/* Copies of the handle's device set arrays (lengths from farm_raster_info);
 * NULL pointers are skipped; layers and lstatus only for -h sets, areas for
 * -h sets and sets farm_rasterize built. */
typedef struct farm_polyset_buf {
    farm_raster_hdr *hdr;
    uint8_t *status;
    uint32_t *pixoff;
    uint32_t *seqs; /* 2 * nseqs */
    uint32_t *voff;
    double *vnodes;
    farm_layer *layers;
    uint8_t *lstatus;
    double *areas;
    uint8_t *fill; /* after farm_set_categories with b != 4 */
} farm_polyset_buf;
int farm_download_polyset(farm_handle h, farm_polyset_buf *out);

/* -w b on the device set (b in 2..8): fill classes at 2^b categories for the
 * approximate mode and the learned features, from the per-pixel clipped
 * areas of a set farm_rasterize built or an -h upload -- the bytes
 * farm_hostset_set_categories gives (encodePixelArea, Ideal.cpp:64-88, with
 * 2^b in place of query_context::category_count, query_context.h:108).
 * b = 4: the status bytes.  Exact results do not change. */
int farm_set_categories(farm_handle h, int category_bits);

/* Upload candidate pairs (ids index the uploaded set); replaces the pair
 * upload at src/cuda/preprocessing.cu:352-358. */
int farm_upload_pairs(farm_handle h, const uint32_t *a, const uint32_t *b, uint64_t n);

/* Candidate pairs generated on the device, replacing the host R-tree probe
 * (rtree_query + MySearchCallback, src/query/intersect_cpu.cpp:7-36 and
 * src/query/within_polygon_cpu.cpp:8-48, RTree::Search
 * src/index/RTree.h:581-600) and the pair upload after it: every (i, j),
 * i < na, j < nb, whose boxes overlap inclusively, box = {low x, low y,
 * high x, high y} as loaded (mbb_a[4i..4i+3], mbb_b[4j..4j+3]); with
 * expand_km > 0 box j is first expanded as box::expand does
 * (Box.cpp:200-214); with skip_self, i == j is skipped.  The pairs become
 * the handle's pair list as (a_base + i, b_base + j), ordered by j then i
 * (the order farm_candidates returns); *npairs gets their count.  Needs the
 * polygons uploaded (a_base + na and b_base + nb within the set) and finite,
 * non-inverted boxes (else FARM_EINVAL); fewer than 2^31 boxes per side and
 * 2^31 pairs (else FARM_ECAPACITY). */
int farm_generate_pairs(farm_handle h, const double *mbb_a, int64_t na, const double *mbb_b, int64_t nb,
                        double expand_km, int skip_self, uint32_t a_base, uint32_t b_base, uint64_t *npairs);

/* Copies pairs [begin, end) of the handle's pair list to host arrays. */
int farm_download_pairs(farm_handle h, uint64_t begin, uint64_t end, uint32_t *a, uint32_t *b);

/* Exact / approximate intersect over pairs [begin, end) (replaces
 * cuda_intersect, src/cuda/query_intersect.cu:335-464, whose batch window is
 * query_context::index/index_end).  out[k - begin] = 1 if pair k intersects.
 * out may be NULL (counts only, as the reference, via stats->found). */
int farm_intersect(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out,
                   farm_stats *stats);

/* Device-pointer variant for callers that keep results resident (bench):
 * out_dev is a device buffer of at least end - begin bytes. */
int farm_intersect_dev(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                       farm_stats *stats);

/* farm_intersect_dev without any device allocation, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No reference counterpart.  The
 * per-call scratch is a fixed device array of the library (about 80 MiB on
 * each GPU the library runs on, held until the process ends): at most
 * 2^18 pairs, 2^20 BORDER x BORDER pixel pairs and 2^20 refinement tiles
 * per call, else FARM_ECAPACITY (out_dev then holds no valid results).
 * Modes: exact batched, exact plain and the reference -a heuristic (the
 * learned modes need the MLP's buffers: FARM_EINVAL).  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from several handles
 * or threads) race; each call starts by clearing what it counts. */
int farm_intersect_noalloc(farm_handle h, uint64_t begin, uint64_t end, const farm_opts *opts, uint8_t *out_dev,
                           farm_stats *stats);

/* Limit on the BORDER x BORDER pixel pairs one farm_intersect call may
 * enumerate (default and maximum 2^32 - 1; 0 restores the default).  A
 * window beyond it returns FARM_ECAPACITY before any refinement runs; split
 * it into smaller batches.  The per-call buffers otherwise grow to fit. */
int farm_set_pixel_pair_limit(farm_handle h, uint64_t max_pixel_pairs);

/* on != 0 (default): the batched intersect replays one captured hipGraph
 * per configuration (mode, tile width, queue order, buffers) for its
 * filter, order and tile launches; 0: plain stream launches. */
int farm_set_graphs(farm_handle h, int on);

/* Exact intersect of pairs [begin, end) that also returns, per pair, the
 * segment tests the batched refinement executed for it (the cost labels of
 * the learned part).  No reference counterpart. */
int farm_intersect_costs(farm_handle h, uint64_t begin, uint64_t end, uint8_t *bits, uint32_t *tests,
                         farm_stats *stats);

/* Within-distance decision over pairs [begin, end) (replaces
 * cuda_within_polygon, src/cuda/query_within_polygon.cu:659-844, with the
 * CPU semantics of Ideal::within, src/geometry/Ideal.cpp:2475-2643): the
 * hierarchical box-distance descent over both polygons' layers, then
 * candidate pixel pairs in dist_apx order refined with the truncated
 * point-segment distances.  Needs a set uploaded with layers and areas.
 * out[k - begin]: 1 within d_km, 0 not, 2 a polygon of the pair is
 * rejected by the reference raster builder.  min_dist (may be NULL): the
 * running minimum segment distance when the decision was taken (100000 when
 * no distance was computed).  out and min_dist may be NULL (results stay on
 * the device).  The device-pointer variant takes device buffers of
 * end - begin entries.  stats: bb_pixel_pairs = candidate pixel pairs of the
 * last level, segment_tests = point-segment distances evaluated,
 * refine_items = box-pair distances of the descent. */
int farm_within(farm_handle h, uint64_t begin, uint64_t end, double d_km, uint8_t *out, double *min_dist,
                farm_stats *stats);
int farm_within_dev(farm_handle h, uint64_t begin, uint64_t end, double d_km, uint8_t *out_dev,
                    double *min_dist_dev, farm_stats *stats);

/* Intersection over pairs [begin, end) of a set rasterized without -h
 * (replaces cuda_intersection, src/cuda/query_intersection.cu:532-754, with
 * the CPU semantics of Ideal::intersection, src/geometry/Ideal.cpp:1440-1683,
 * which stops after classifying the boundary pieces).  counts (5 per pair,
 * may be NULL): crossing points after de-duplication (-1: every source pixel
 * over the target MBB is OUT, -2: every one IN, -3: a polygon is rejected by
 * the raster builder), boundary pieces, pieces IN / OUT / on the boundary of
 * the other polygon.  area (may be NULL): area of A n B, integrated over the
 * boundary pieces IN the other polygon (Green's theorem, f64); no reference
 * counterpart.  stats: bb_pixel_pairs, segment_tests (edge pairs tested),
 * refine_items (crossing records before de-duplication), found (pairs with
 * area > 0).  A pair may have up to 2^26 crossing records (else
 * FARM_ECAPACITY). */
int farm_intersection_area(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts, double *area,
                           farm_stats *stats);
int farm_intersection_area_dev(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts_dev, double *area_dev,
                               farm_stats *stats);

/* One piece of the boundary of A n B: a run of one polygon's boundary that
 * lies IN the other polygon, from crossing point a through the ring
 * vertices first .. first + nvert - 1 (indexes into the polygon's decoded
 * ring, taken modulo its m = nverts - 1 distinct vertices) to crossing point
 * b.  nvert == -1: the whole ring (one polygon contains the other).  The
 * pieces of a pair chain into the rings of A n B (each piece's b is the a of
 * the next), the assembly the reference's rebuild_polygons performs
 * (src/cuda/query_intersection.cu:393-502, never launched there). */
typedef struct farm_piece {
    uint32_t pair;  /* index in the call's pair window */
    int32_t ring;   /* 0: the source polygon, 1: the target polygon */
    int32_t first;
    int32_t nvert;
    float ax, ay, bx, by;
} farm_piece;

/* farm_intersection_area that also returns the IN pieces of every pair with
 * an intersection (in no particular order; farm_pieces_wkt assembles them).
 * *npieces gets their count; with more than cap of them nothing is copied
 * and the call returns FARM_ECAPACITY (repeat with cap >= *npieces).  counts
 * and area may be NULL. */
int farm_intersection_pieces(farm_handle h, uint64_t begin, uint64_t end, int32_t *counts, double *area,
                             farm_piece *pieces, uint64_t cap, uint64_t *npieces, farm_stats *stats);

/* ---- learned part (no reference counterpart) ----
 * Pair features: FARM_NFEAT floats per candidate pair from the filter's
 * pixel windows: 64-bin histogram of BORDER x BORDER fill-ratio products,
 * log counts of IN and BORDER pixel pairs, log vertex counts, MBR overlap
 * ratio, log refinement cost bound, heuristic probabilities, log edge sums,
 * pixel-area ratio, window BORDER / IN fractions, zero padding.
 * decided[k]: 1 an IN overlap (exact answer true), 0 no BORDER pixel pair
 * (exact answer false), -1 needs refinement. */
#define FARM_NFEAT 80
#define FARM_MLP_NPARAMS 27138 /* 80x128 + 128 + 128x128 + 128 + 128x2 + 2 */
int farm_pair_features(farm_handle h, uint64_t begin, uint64_t end, float *feat, int8_t *decided);
/* MLP 80 -> 128 (ReLU) -> 128 (ReLU) -> 2 (intersect logit, log2(1 + tests)).
 * params: f32 master weights [W1 (80x128), b1, W2 (128x128), b2, W3 (128x2),
 * b3], row-major; norm: per-feature mean then 1/std (2 x 80) applied to raw
 * features on the device (NULL = identity).  Products run on MFMA with bf16
 * operands and f32 accumulation. */
int farm_mlp_set_params(farm_handle h, const float *params, uint64_t n, const float *norm);
int farm_mlp_get_params(farm_handle h, float *params, uint64_t n);
/* X: n x 80 normalised features (host); out: n x 2 */
int farm_mlp_forward(farm_handle h, const float *X, uint64_t n, float *out);
/* loss = mean BCE(logit, y) + lambda * mean (cost - c)^2 and its gradient
 * (FARM_MLP_NPARAMS floats) on one batch */
int farm_mlp_loss_grad(farm_handle h, const float *X, const float *y, const float *c, uint64_t n, float lambda,
                       float *loss, float *grad);
/* Adam training on the device (shuffled minibatches); epoch_loss[epochs] */
int farm_mlp_train(farm_handle h, const float *X, const float *y, const float *c, uint64_t n, int epochs, int batch,
                   float lr, float lambda, uint64_t seed, float *epoch_loss);
/* features + normalisation + forward for pairs [begin, end); out: n x 2 */
int farm_mlp_predict_pairs(farm_handle h, uint64_t begin, uint64_t end, float *out);

/* Launch configuration of the batched refinement. */
typedef struct farm_tuning {
    int32_t tile_a;        /* source edges per lane: 1, 2, 4 or 8 */
    int32_t chunk;         /* target segments per chunk box: 8, 16 or 32 */
    int32_t blocks_per_cu; /* persistent refinement workgroups per CU: 1..32 */
    int32_t shape_bucket;  /* output: the shape bucket of the uploaded set this entry is for
                              (half-octaves of its mean edges per BORDER pixel) */
    double ms;             /* output: device time of the tuning window with it (0: not timed) */
} farm_tuning;

/* Autotuner of the batched refinement: times pairs [begin, end) with every
 * tile width (1, 2, 4, 8) and chunk size (8, 16, 32), then with 4, 8 and 16
 * workgroups per CU at the fastest pair, and keeps the fastest as the
 * configuration of later calls whose opts->tile_a / tile_b are 0 and as the
 * entry of the set's shape bucket (see farm_tuning_save).  No reference
 * counterpart (the reference hard-codes 16x16 tiles,
 * src/cuda/query_intersect.cu:194). */
int farm_autotune_config(farm_handle h, uint64_t begin, uint64_t end, farm_tuning *best);
/* farm_autotune_config, reporting only the tile width. */
int farm_autotune(farm_handle h, uint64_t begin, uint64_t end, int *best_tile);
/* Set the tile width used when opts->tile_a is 0 (e.g. a persisted choice). */
int farm_set_tile(farm_handle h, int tile);
/* The configuration calls with opts->tile_a / tile_b == 0 use, and the set's
 * shape bucket. */
int farm_get_tuning(farm_handle h, farm_tuning *out);
/* Make t that configuration and the entry of the set's shape bucket
 * (rebuilds the chunk boxes when t->chunk differs from the built size). */
int farm_set_tuning(farm_handle h, const farm_tuning *t);
/* Persist the per-bucket entries as text ("farm-tuning 1", then one line
 * "bucket tile_a chunk blocks_per_cu ms" per entry), and read them back:
 * load merges the file's entries and applies the uploaded set's bucket when
 * the file holds it.  FARM_EINVAL for an unreadable or malformed file. */
int farm_tuning_save(farm_handle h, const char *path);
int farm_tuning_load(farm_handle h, const char *path);

/* HBM calibration: times `iters` launches of a plain streaming copy of
 * `bytes` bytes between two scratch buffers; *gbps = bytes read + written
 * per second.  Its byte counters are the reference point for the traffic
 * figures of the other kernels.  No reference counterpart. */
int farm_copy_probe(farm_handle h, uint64_t bytes, int iters, double *gbps);

/* Library version string. */
const char *farm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FARM_CAPI_H */
