/*
 * farm_oracle.h -- CPU restatement of the reference FARM CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (farm_amd/) links,
 * includes or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / baseline.
 *
 * The restatement follows the reference sources function by function; each
 * function in farm_oracle.cpp cites the reference file:line it restates.
 *
 * Parity pinning: the reference publishes no golden vectors and cannot be
 * built in this image (it needs boost headers that are absent), so every
 * function here is pinned by tests/test_oracle_*.py against closed forms,
 * invariants and brute force (see the header of farm_oracle.cpp for the
 * function -> check map).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes returned by every orc_* entry point. */
#define ORC_OK 0
#define ORC_EINVAL 1
/* the reference itself would abort (an assert fires) on this input */
#define ORC_EREF_ABORT 2

typedef struct orc_set orc_set;

/* Build a polygon set from rings exactly as stored in a .dt file (the ring
 * includes the WKT closing vertex; decode appends p[0] once more).
 * ring_off has n+1 entries indexing xy pairs. Rings with fewer than 2 stored
 * vertices are dropped, as the loader does. */
orc_set *orc_create(int64_t n, const uint64_t *ring_off, const float *xy);
void orc_destroy(orc_set *s);
int64_t orc_size(const orc_set *s);

/* Rasterize every polygon: init_raster(nv / vpr) then Ideal::rasterization.
 * hierarchy != 0 additionally runs grid_align + layering(nlow) and merges
 * layer status (the -h path).  Returns ORC_OK or ORC_EREF_ABORT (index of
 * the first offending polygon in *bad). */
int orc_rasterize(orc_set *s, int vpr, int hierarchy, int nlow, int nthreads, int64_t *bad);

/* Raster accessors for parity tests. */
int orc_raster_info(const orc_set *s, int64_t i, double *mbr4, int *dims2, double *steps2,
                    int64_t *counts /* [nstatus, nseqs, nvnodes, nhnodes, nverts] */);
int orc_raster_copy(const orc_set *s, int64_t i, uint8_t *status, uint32_t *offset,
                    uint32_t *seqs /* 2 per seq */, uint32_t *voff, double *vnodes,
                    uint32_t *hoff, double *hnodes, double *areas);

/* Layer k (0 = coarsest) of polygon i after orc_rasterize(..., hierarchy=1):
 * grid origin, square step, dims and (if status != NULL) its fill classes.
 * Returns the number of layers (k out of range: nothing written), -1 for a
 * bad polygon index. */
int orc_layer(const orc_set *s, int64_t i, int k, double *low2, double *step, int *dims2, uint8_t *status);

/* Exact intersect over candidate pairs (source set a, target set b):
 * out[k] = Ideal::intersect(a[pa[k]], b[pb[k]]).  mode 0 = exact,
 * mode 1 = reference -a heuristic (IN overlap or a BORDER pixel pair with
 * probability > 1).  nthreads workers take batches of 10 pairs.
 * stats (may be NULL): [pixel_pairs_bb, segment_tests, in_hits]. */
int orc_intersect(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa,
                  const uint32_t *pb, uint8_t *out, int mode, int nthreads, int64_t *stats);

/* Ideal::within per ordered pair (source a[pa[k]], target b[pb[k]]) with the
 * reference's distance convention (km); needs orc_rasterize(..., hierarchy=1).
 * out: 1 within, 0 not, 2 the reference aborts; min_dist (may be NULL): the
 * running minimum segment distance of the refinement (100000 if none ran).
 * stats (may be NULL): [candidate pixel pairs, point-segment distances]. */
int orc_within(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa, const uint32_t *pb,
               double d_km, uint8_t *out, double *min_dist, int nthreads, int64_t *stats);

/* Ideal::intersection per pair (source a[pa[k]], target b[pb[k]]) up to the
 * reference's boundary-piece statuses, plus the intersection area by Green's
 * theorem over the pieces (see farm_oracle.cpp).  counts (5 per pair, may be
 * NULL): [intersection points after dedup (-1 window all OUT, -2 window all
 * IN, -3 the reference aborts), pieces, IN, OUT, BORDER]; area (may be
 * NULL): f64 area of A n B. */
int orc_intersection(const orc_set *a, const orc_set *b, int64_t npairs, const uint32_t *pa, const uint32_t *pb,
                     int32_t *counts, double *area, int nthreads);

/* Within self-join candidates: (i, j), i != j, with i's MBB overlapping j's
 * MBB expanded by d_km (box::expand). */
int64_t orc_candidates_within(const orc_set *s, double d_km, int64_t cap, uint32_t *pa, uint32_t *pb);

/* Candidate generation by brute force: all (i, j) with inclusively
 * overlapping MBBs; writes up to cap pairs, returns the total count. */
int64_t orc_candidates(const orc_set *a, const orc_set *b, int64_t cap, uint32_t *pa, uint32_t *pb);

/* MBB of polygon i (low x, low y, high x, high y) as the loader computes it. */
int orc_mbb(const orc_set *s, int64_t i, double *mbr4);

/* 1 if the reference raster builder aborts (or never ends) on polygon i. */
int orc_failed(const orc_set *s, int64_t i);

/* Primitive restatements exposed for unit tests. */
int orc_segment_intersect(const float *abcd /* ax ay bx by cx cy dx dy */);
int orc_double_to_int(double v);
int64_t orc_retrieve(const orc_set *s, int64_t i, const double *box4, int unshrunk, int32_t *ids, int64_t cap);
int orc_pixel_box(const orc_set *s, int64_t i, int x, int y, double *out4);
double orc_point_to_segment_distance(const float *p, const float *p1, const float *p2, int geography);

#ifdef __cplusplus
}
#endif
