/*
 * oracle.h -- CPU restatement of the reference's SFC encode -> reorder -> linear-octree build ->
 * radius / kNN search path (derivada/octrees-benchmark).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the timed CPU
 * baseline). The GPU library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * repository root). Arithmetic follows the reference's operation order with no FP contraction
 * (build flag -ffp-contract=off), because the reference's x86 Release build (-O2 -mavx2 -mbmi2,
 * cmake/CMakeConfig.cmake:41,44) has no FMA.
 *
 * Parity pinning (what checks each function against something other than itself):
 *   orc_mbb                : reference KATs (test_geometry.cpp:38-51, test_encoders.cpp:46-50) + closed form
 *   orc_quantize/encode    : closed form incl. the x86 double->uint64 edge cases (measured with g++)
 *   orc_morton_*           : vendored libmorton compiled from the reference sources (oracle/_ref),
 *                            reference KAT test_complex.cpp:96-114
 *   orc_hilbert_*          : invariants only (bijection, unit-step adjacency, start cell, level-1 table
 *                            hilbert_encoder_3d.hpp:22); the exact orientation convention is
 *                            "parity unpinned" beyond these invariants (the reference needs Eigen and
 *                            PAPI headers that this image lacks, so it cannot be compiled here)
 *   orc_sort               : closed form (= stable sort by key), reference KAT test_encoders.cpp:36-44
 *   orc_build              : closed form (canonical "split iff count > maxLeaf" tree, BFS order) and
 *                            brute-force recursion on small inputs
 *   orc_radius             : brute force over all points with the same predicate (equal except in
 *                            rounding-pathological node classifications), reference KATs
 *                            test_octree.cpp:63-80, test_octree_advanced.cpp:69-131,
 *                            test_complex.cpp:37-94
 *   orc_knn / orc_knn_brute: brute-force (d2, index) order; reference KATs test_octree.cpp:42-61,
 *                            test_octree_advanced.cpp:33-67, test_complex.cpp:144-172
 *   orc_las_*              : LAS files written field by field from the ASPRS LAS 1.2 / 1.3 / 1.4
 *                            layouts by an independent writer (tests/las_files.py) with known
 *                            integers and attributes, and power-of-two scales whose coordinates have a
 *                            closed form. The reference reads through LASlib (LAStools, cloned at its
 *                            head by scripts/install_laslib.sh, so no version is pinned), absent here:
 *                            the attributes of point formats 6-10 follow LASlib's legacy view as
 *                            restated in orc_las_read and are "parity unpinned"; formats 0-5 and all
 *                            coordinates are pinned by the files' own fields.
 */
#ifndef OCTREE_ORACLE_H
#define OCTREE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Encoder ids: values of the reference's EncoderType enum (inc/main_options.hpp:73-83). */
#define ORC_ENC_MORTON3D 1
#define ORC_ENC_HILBERT3D 5
/* Kernel ids: values of the reference's Kernel_t enum (inc/main_options.hpp:97-103). */
#define ORC_K_CIRCLE 0
#define ORC_K_SPHERE 1
#define ORC_K_SQUARE 2
#define ORC_K_CUBE 3
/* Search algorithms: values of the reference's SearchAlgo enum (inc/main_options.hpp:29-35). */
#define ORC_A_NEIGHBORS 1
#define ORC_A_PRUNE 2
#define ORC_A_STRUCT 3
#define ORC_A_APPROX 4 /* neighborsApprox: Struct with a depth cutoff */

/* geometry (inc/geometry/box.hpp:94-137, util.hpp:139-143) */
void orc_mbb(const double *xyz, size_t n, double center[3], double radii[3]);
/* quantisation (point_encoder.hpp:37-49) with the x86 conversion semantics */
uint64_t orc_quantize_axis(double p, double c, double r);
uint64_t orc_morton_encode(uint64_t x, uint64_t y, uint64_t z);
void orc_morton_decode(uint64_t key, uint64_t out[3]);
uint64_t orc_hilbert_encode(uint64_t x, uint64_t y, uint64_t z);
void orc_hilbert_decode(uint64_t key, uint64_t out[3]);
/* encodePoints (point_encoder.hpp:54-63) */
void orc_encode(int enc, const double *xyz, size_t n, const double center[3], const double radii[3],
                uint64_t *keys);
/* sortPoints radix passes (point_encoder.hpp:285-383): keys sorted in place, perm[i] = original
 * index of the point now at position i. nthreads emulates the per-thread histogram layout. */
void orc_sort(uint64_t *keys, uint32_t *perm, size_t n, int nthreads);
/* out[i] = xyz[perm[i]] (n x 3 doubles): the points moved with their keys (point_encoder.hpp:359-375). */
void orc_gather3(const double *xyz, const uint32_t *perm, size_t n, double *out);

/* ---------------- linear octree (inc/structures/linear_octree.hpp) ---------------- */
typedef struct orc_tree orc_tree;
/* Builds from sorted codes and the sortPoints Box(center, radii). Returns NULL on failure. */
orc_tree *orc_build(const uint64_t *codes, size_t n, uint32_t max_points_leaf, int enc,
                    const double center[3], const double radii[3]);
void orc_tree_free(orc_tree *t);
/* counts: nLeaf, nInternal, nTotal, maxDepthSeen, rounds */
void orc_tree_info(const orc_tree *t, uint32_t info[5]);
const uint64_t *orc_tree_leaves(const orc_tree *t);   /* nLeaf+1 */
const uint64_t *orc_tree_counts(const orc_tree *t);   /* nLeaf */
const uint64_t *orc_tree_layout(const orc_tree *t);   /* nLeaf+1 */
const uint64_t *orc_tree_prefixes(const orc_tree *t); /* nTotal */
const uint32_t *orc_tree_offsets(const orc_tree *t);  /* nTotal+1 */
const uint64_t *orc_tree_ranges(const orc_tree *t);   /* nTotal x 2 */
const double *orc_tree_centers(const orc_tree *t);    /* nTotal x 3 */
const double *orc_tree_radii(const orc_tree *t);      /* 22 x 3 */

/* ---------------- searches ---------------- */
typedef struct orc_result orc_result;
/* One search per query. centers: nq x 3 doubles; radii: nq doubles. Output per query in the
 * reference's own emission order: neighbors/Prune -> point indices; Struct -> [first,last) ranges
 * (including empty INSIDE ranges, neighbor_set.hpp:84-89). stats (nullable, nq x 2): per query
 * (points tested with isInside, nodes classified). */
orc_result *orc_radius(const orc_tree *t, const double *sorted_xyz, int kernel, int algo,
                       const double *centers, const double *radii, size_t nq, uint64_t *stats);
/* Cube / square searches with per-axis half extents (nq x 3), Vector overloads of
 * linear_octree.hpp:1044-1060, 1069-1075. */
orc_result *orc_radius_box(const orc_tree *t, const double *sorted_xyz, int kernel, int algo, const double *centers,
                           const double *radii3, size_t nq);
/* neighborsApprox (linear_octree.hpp:981-1035): Struct with a depth cutoff; a node deeper than
 * `level` is added whole (upper != 0) or dropped, without a kernel test. radii3 NULL: scalar radii. */
orc_result *orc_radius_approx(const orc_tree *t, const double *xyz, int kernel, const double *centers,
                              const double *radii, const double *radii3, size_t nq, uint32_t level, int upper);
size_t orc_result_total(const orc_result *r);        /* total entries (indices, or ranges) */
const uint64_t *orc_result_offsets(const orc_result *r); /* nq+1 offsets into entries */
const uint64_t *orc_result_entries(const orc_result *r); /* indices, or 2*ranges interleaved */
const uint64_t *orc_result_sizes(const orc_result *r);   /* nq: number of neighbour points */
void orc_result_free(orc_result *r);
/* Count-only variant (what the reference harness keeps: result.size()); returns sum of sizes. */
uint64_t orc_radius_count(const orc_tree *t, const double *sorted_xyz, int kernel, int algo,
                          const double *centers, const double *radii, size_t nq, uint64_t *sizes);
/* The reference harness' per-query work: build each result, keep its size (timed CPU baseline). */
uint64_t orc_radius_harness(const orc_tree *t, const double *sorted_xyz, int kernel, int algo,
                            const double *centers, const double *radii, size_t nq);
/* Brute force with the same per-point predicate (test aid). */
orc_result *orc_radius_brute(const double *sorted_xyz, size_t n, int kernel, const double *centers,
                             const double *radii, size_t nq);

/* kNN best-first (linear_octree.hpp:819-942). idx/d2: nq x k, unused slots = -1 / NaN.
 * Returns total found. */
uint64_t orc_knn(const orc_tree *t, const double *sorted_xyz, size_t n, const double *centers,
                 size_t nq, uint32_t k, int64_t *idx, double *d2);
/* Brute force k smallest under the (d2, index) order. */
uint64_t orc_knn_brute(const double *sorted_xyz, size_t n, const double *centers, size_t nq,
                       uint32_t k, int64_t *idx, double *d2);

int orc_max_threads(void);

/* Competitor baseline: the reference's pointer-based Octree (inc/structures/octree.hpp:20-396,
 * src/octree.cpp:17-322) over xyz (n x 3, kept by reference), insertion-built from the root box
 * (center, radii); orc_ptree_info: leaves, internal nodes, depth. orc_ptree_radius: its neighbors()
 * search (boxOverlap descent, isInside per leaf point) for nq centres of one radius, OpenMP over the
 * centres; sizes[q] per centre; with out_idx and offsets (exclusive prefix of sizes) the index lists
 * in the reference's emission order. Returns the total. */
typedef struct orc_ptree orc_ptree;
orc_ptree *orc_ptree_build(const double *xyz, size_t n, uint32_t max_points_leaf, const double center[3],
                           const double radii[3]);
void orc_ptree_free(orc_ptree *t);
void orc_ptree_info(const orc_ptree *t, uint64_t info[3]);
uint64_t orc_ptree_radius(const orc_ptree *t, int kernel, const double *centers, size_t nq, double radius,
                          uint64_t *sizes, const uint64_t *offsets, uint64_t *out_idx);

/* LAS ingest as the reference's reader sees it (readers/las_file_reader_parallel.hpp:52-64, 74-95,
 * 112-137 through LASlib). orc_las_header: info[0] version minor, [1] point format, [2] record length,
 * [3] number of points (legacy count, or the LAS 1.4 64-bit count when the legacy one is 0),
 * [4] offset to the point records; so[6] = x, y, z scale then x, y, z offset. Returns 0, or -1 for
 * an unreadable file / bad signature, -2 for LAZ or an undefined format.
 * orc_las_read: xyz (n x 3) = X * scale + offset (:88-92); attributes as PointMetadata
 * (:124-137): intensity (n), fields (n x 8: return number, number of returns, scan direction, edge
 * of flight line, classification, scan angle (as int8), user data, 0), point source id (n), rgb (n x 3). */
int orc_las_header(const char *path, uint64_t info[5], double so[6]);
int orc_las_read(const char *path, double *xyz, double *intensity, uint8_t *fields, uint16_t *psid, uint32_t *rgb);

#ifdef __cplusplus
}
#endif
#endif
