/*
 * farm_host.h -- host side of the MI355X polygon-join library (libfarm.so):
 * .dt / WKT polygon storage, candidate-pair generation (host R-tree) and the
 * host raster builder that produces the device layout of farm_capi.h.
 *
 * These replace the reference's host preprocessing:
 *   farm_hostset_load_dt      load_binary_file   src/storage/storage.cpp:361-449
 *   farm_hostset_save_dt      dump_polygons_to_file  src/storage/storage.cpp:158-193
 *   farm_hostset_load_wkt     load_polygon_wkt   src/storage/storage.cpp:45-94
 *   farm_hostset_rasterize    preprocess / process_rasterization
 *                             src/geometry/preprocessing.cpp:4-87 (+ Ideal.cpp:920-955)
 *   farm_candidates           rtree_query + MySearchCallback
 *                             src/query/intersect_cpu.cpp:7-36, within_polygon_cpu.cpp:8-48
 */
#ifndef FARM_HOST_H
#define FARM_HOST_H

#include "farm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct farm_hostset farm_hostset;

/* Build from rings as stored in .dt files (closing vertex included; the
 * loader appends the first vertex once more, vertex_sequence.cpp:206-218).
 * Rings with fewer than 2 stored vertices are dropped (storage.cpp:341). */
farm_hostset *farm_hostset_from_rings(int64_t n, const uint64_t *ring_off, const float *xy);
/* Ids follow file order (the reference's multi-threaded loader may permute
 * files larger than 10 MiB, storage.cpp:355-357,447). */
farm_hostset *farm_hostset_load_dt(const char *path);
/* One polygon per line: "POLYGON ((x y, ...), (hole), ...)"; coordinates are
 * parsed to float as std::stof does (util.h:129-144).  Holes are kept in the
 * file but not used by queries (MyPolygon.cpp:754-757). */
farm_hostset *farm_hostset_load_wkt(const char *path);
int farm_hostset_save_dt(const farm_hostset *s, const char *path);
/* Concatenate (sources then targets) before rasterizing. */
farm_hostset *farm_hostset_concat(const farm_hostset *a, const farm_hostset *b);
void farm_hostset_free(farm_hostset *s);
int64_t farm_hostset_size(const farm_hostset *s);
/* Stored ring of polygon i (without the appended copy of vertex 0). */
int farm_hostset_ring(const farm_hostset *s, int64_t i, const float **xy, int64_t *nv);
/* MBB (low x, low y, high x, high y) with the float 1e-9 bias of
 * VertexSequence::getMBR (vertex_sequence.cpp:152-167). */
int farm_hostset_mbbs(const farm_hostset *s, double *out4n);
/* farm_rasterize for a host set, with the rings streamed to the device
 * through two pinned staging buffers (host threads convert chunk k + 1 into
 * closed SoA rings while chunk k's DMA runs): the same set, without the
 * full-size host copy of farm_hostset_closed_rings or a pageable upload.
 * Replaces the upload half of preprocess_for_gpu (src/cuda/preprocessing.cu:
 * 128-343) together with the rasterization (see farm_rasterize). */
int farm_hostset_rasterize_device(farm_handle h, const farm_hostset *s, int vpr, int nlow, int nthreads,
                                  farm_raster_info *info);
/* Vertices per pinned staging buffer of farm_hostset_rasterize_device
 * (default 2^22, 1,024 .. 2^28; two buffers of 8 B per vertex); a ring
 * longer than that goes through a pinned buffer of its own.  No counterpart
 * in the reference (its upload is one pageable copy). */
int farm_set_staging(farm_handle h, uint64_t vertices);
/* The rings as the raster builder reads them (each stored ring plus its
 * first vertex once more), for farm_rasterize: vert_off gets npolys + 1
 * offsets; x / y (may be NULL: offsets only) vert_off[npolys] floats each. */
int farm_hostset_closed_rings(const farm_hostset *s, uint64_t *vert_off, float *x, float *y, int nthreads);

/* Rasterize every polygon with vpr vertices per pixel (-v), nthreads host
 * threads (-n).  Fills the device layout returned by farm_hostset_view. */
int farm_hostset_rasterize(farm_hostset *s, int vpr, int nthreads);
/* -h: the raster on the grid snapped to the universal steps plus the coarse
 * layers down to at most nlow pixels (--NLow), with per-pixel areas
 * (MyRaster::grid_align, MyRaster.cpp:444-491; Ideal::layering and
 * merge_status, Ideal.cpp:1008-1099).  Needed by farm_within. */
int farm_hostset_rasterize_h(farm_hostset *s, int vpr, int nlow, int nthreads);
/* Layer k (0 = coarsest, nlayers - 1 = the raster itself) of polygon i of a
 * set rasterized with farm_hostset_rasterize_h, for tests. */
int farm_hostset_layer(const farm_hostset *s, int64_t i, int k, farm_layer *out, const uint8_t **status);
int farm_hostset_view(const farm_hostset *s, farm_polyset *out);
/* -w b: fill classes of every pixel at 2^b categories (b in 2..8), from the
 * same clipped areas (encodePixelArea, Ideal.cpp:64-88, with 2^b in place of
 * 16), for the approximate mode and the learned features; b = 4 keeps the
 * status bytes alone.  Call after rasterizing; farm_hostset_view then
 * exposes them. */
int farm_hostset_set_categories(farm_hostset *s, int category_bits);
/* Per-pixel clipped areas of polygon i (dimx*dimy doubles), for tests. */
int farm_hostset_areas(const farm_hostset *s, int64_t i, const double **areas);
/* Index of the first polygon the reference raster builder would reject
 * (an assert), or -1. */
int64_t farm_hostset_raster_failures(const farm_hostset *s);
/* Per polygon: 1 if rasterized, 0 if the reference raster builder would
 * abort or never terminate on it (the polygon is then kept as an empty
 * raster: every pixel OUT, so every query on it answers false). */
int farm_hostset_raster_ok(const farm_hostset *s, uint8_t *out);

/* Candidate pairs: every (i in a, j in b) whose MBBs overlap inclusively
 * (RTree::Overlap, src/index/RTree.h:1567-1580).  expand_km > 0 expands b's
 * MBB with box::expand (Box.cpp:200-214); skip_self drops i == j (within
 * self-join).  Pairs come out sorted by (j, i).  Arrays are allocated by the
 * library; free them with farm_free. */
int64_t farm_candidates(const farm_hostset *a, const farm_hostset *b, double expand_km, int skip_self,
                        int nthreads, uint32_t **pa, uint32_t **pb);
void farm_free(void *p);

/* Seeded synthetic star-shaped polygons, generated on nthreads host threads
 * (the same set for any thread count): n closed rings, centres uniform in
 * [lo, hi]^2, vertex counts uniform in [nv_lo, nv_hi], sorted uniform
 * angles, radius R (1 + amp u) per vertex with u ~ U(-1, 1) and
 * R ~ LogNormal(r_mu, r_sigma), coordinates rounded to float32.  With
 * nv 50..500, R ~ LogNormal(ln 0.5, 0.5), amp 0.4 and a [0, 100]^2 box per
 * 2,000 polygons this is the plain intersect workload's distribution.
 * Returns NULL on bad arguments. */
farm_hostset *farm_hostset_star_polygons(uint64_t seed, int64_t n, double lo, double hi, int nv_lo, int nv_hi,
                                         double r_mu, double r_sigma, double amp, int nthreads);
/* The same generator with the other workloads' options: pareto_alpha > 0
 * draws vertex counts from Pareto(alpha) with scale nv_lo, floored and
 * clipped to [nv_lo, nv_hi] (complex polygons: 1.2, 1,000, 200,000);
 * noise 1 replaces the per-vertex noise by three octaves of radial noise,
 * r = R (1 + amp s / 1.75) with s = sum_o 0.5^o sin(f 2^o theta + phi_o),
 * f ~ U{2..6} and phi_o ~ U[0, 2 pi) per polygon. */
farm_hostset *farm_hostset_star_polygons_ex(uint64_t seed, int64_t n, double lo, double hi, int nv_lo, int nv_hi,
                                            double pareto_alpha, double r_mu, double r_sigma, int noise, double amp,
                                            int nthreads);

/* The polygons of A n B assembled from farm_intersection_pieces output, as
 * WKT: one line "<k>\tMULTIPOLYGON (((x y, ...)), ...)" per pair k of the
 * window that has pieces, k ascending.  s is the rasterized set that was
 * uploaded, pa / pb the window's pair ids into it.  Pieces chain by their
 * crossing points (each piece's b is the next one's a, compared as floats);
 * a chain that does not close is closed at its start.  Returns text to free
 * with farm_free (*len its length), NULL on bad input. */
char *farm_pieces_wkt(const farm_hostset *s, const uint32_t *pa, const uint32_t *pb, uint64_t npairs,
                      const struct farm_piece *pieces, uint64_t npieces, uint64_t *len);

/* The 900-entry degrees-per-kilometre longitude table the library uses for
 * within distances and the within join's box expansion: out[k] covers
 * latitudes [k/10, (k+1)/10) (degree_per_kilometer_longitude_arr,
 * src/include/util.h:73-74). */
int farm_km_table(double *out);

#ifdef __cplusplus
}
#endif

#endif /* FARM_HOST_H */
