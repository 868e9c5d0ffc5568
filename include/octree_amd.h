/*
 * octree_amd.h -- C ABI of the MI355X (gfx950) SFC-encode / reorder / linear-octree / neighbour-
 * search library. Plain C, extern "C", plain pointers and sizes; no C++ or torch types cross it.
 *
 * Each entry point replaces one piece of the reference's header-only C++ API
 * (derivada/octrees-benchmark; paths relative to its root). The drop-in C++ facade
 * (headers under octrees-benchmark_amd/facade/, laid out as the reference's inc/) keeps the reference's class names and signatures and
 * forwards to these calls.
 *
 * Conventions
 *  - Every call returns OCT_OK (0) or a negative OCT_ERR_* code; oct_last_error() gives the text.
 *    Nothing throws across the ABI and nothing calls exit() (the reference exit()s on overflow at
 *    linear_octree.hpp:387-390; here that is OCT_ERR_STATE).
 *  - A context owns one HIP stream on one device and every device buffer; calls on one context are
 *    not re-entrant (the facade serialises them with a mutex).
 *  - Host buffers are copied in/out; "_dev" variants take/return device pointers on the context's
 *    device and enqueue on the context's stream.
 *  - Indices are positions in the SFC-sorted cloud, as in the reference (searches run after
 *    sortPoints, so the reference returns sorted-cloud indices).
 */
#ifndef OCTREE_AMD_H
#define OCTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oct_ctx oct_ctx;

enum oct_status {
    OCT_OK = 0,
    OCT_ERR_ARG = -1,         /* invalid argument (null pointer, unknown id, size limit) */
    OCT_ERR_HIP = -2,         /* HIP runtime error */
    OCT_ERR_STATE = -3,       /* call out of order (e.g. search before build) or device check failed */
    OCT_ERR_NOMEM = -4,       /* device allocation failed */
    OCT_ERR_UNSUPPORTED = -5  /* encoder/kernel combination the reference does not support */
};

/* Encoder ids = the reference's EncoderType values (inc/main_options.hpp:73-83). Only the two 3-D
 * encoders can build a LinearOctree (inc/benchmarking/neighbor_benchmarks.hpp:782-786). */
enum oct_encoder { OCT_ENC_MORTON3D = 1, OCT_ENC_HILBERT3D = 5 };
/* Kernel ids = the reference's Kernel_t values (inc/main_options.hpp:97-103). */
enum oct_kernel { OCT_KERNEL_CIRCLE = 0, OCT_KERNEL_SPHERE = 1, OCT_KERNEL_SQUARE = 2, OCT_KERNEL_CUBE = 3 };
/* Search algorithms = the reference's SearchAlgo values (inc/main_options.hpp:29-35). */
enum oct_algo { OCT_ALGO_NEIGHBORS = 1, OCT_ALGO_PRUNE = 2, OCT_ALGO_STRUCT = 3, OCT_ALGO_APPROX = 4 /* oct_radius_approx */ };
/* Arithmetic of the per-point / per-node tests. FP64 evaluates the reference's expressions as
 * written. FP32_RECHECK evaluates them in FP32 on leaf-recentred coordinates and re-evaluates in
 * FP64 every decision that falls inside a proven rounding band; results are identical. */
enum oct_precision { OCT_PREC_FP64 = 0, OCT_PREC_FP32_RECHECK = 1 };

typedef struct oct_tree_info {
    uint32_t n_leaf;          /* leaves (cornerstone nLeaf) */
    uint32_t n_internal;      /* internal nodes, (nLeaf - 1) / 7 */
    uint32_t n_total;         /* nLeaf + nInternal */
    uint32_t max_depth_seen;  /* deepest level created by a split */
    uint32_t max_points_leaf; /* build parameter */
    uint32_t n_levels;        /* levels holding nodes */
    double build_ms;          /* device time of the build kernels */
} oct_tree_info;

typedef struct oct_search_stats {
    uint64_t total_neighbors;  /* sum over queries of |N(q)| */
    uint64_t points_tested;    /* per-point predicate evaluations */
    uint64_t nodes_classified; /* node predicate evaluations */
    uint64_t rechecked;        /* FP32 decisions re-evaluated in FP64 (points + nodes) */
    double kernel_ms;          /* device time of the search kernel(s) */
    uint64_t class_slots;      /* wave lane-slots spent on node tests (64 per wave-level test);
                                  kNN: wave-wide candidate-buffer selections */
    uint64_t test_slots;       /* wave lane-slots spent on point tests; kNN: wave-wide leaf steps */
} oct_search_stats;

/* ---- context -------------------------------------------------------------------------------- */
int oct_create(oct_ctx **ctx, int device);
int oct_destroy(oct_ctx *ctx);
const char *oct_last_error(const oct_ctx *ctx);
int oct_synchronize(oct_ctx *ctx);
const char *oct_version(void);

/* ---- H1 points + bounding box ----------------------------------------------------------------
 * oct_set_points: copies n points; stride_bytes is the byte distance between consecutive points
 * whose first three doubles are x, y, z (24 for packed xyz, 32 for the reference's Point struct,
 * inc/geometry/point.hpp:23-29). Replaces the container hand-off of sortPoints.
 * oct_bbox: mbb (inc/geometry/box.hpp:94-137) -> center = min + (max - min) / 2,
 * radii = (max - min) / 2; the encode/octree Box is Box(center, radii) (box.hpp:17). */
int oct_set_points(oct_ctx *ctx, const double *xyz, size_t n, size_t stride_bytes);
int oct_set_points_dev(oct_ctx *ctx, const double *xyz_dev, size_t n, size_t stride_bytes);
/* The structure-of-arrays container PointsSoA (inc/geometry/point_containers.hpp:72-163): n points
 * as three host coordinate arrays (its dataX / dataY / dataZ). Same effect as oct_set_points. */
int oct_set_points_soa(oct_ctx *ctx, const double *x, const double *y, const double *z, size_t n);
int oct_bbox(oct_ctx *ctx, double center[3], double radii[3]);

/* ---- H2 encode: PointEncoder::encodePoints (point_encoder.hpp:54-63) with getAnchorCoords
 * (:37-49) and MortonEncoder3D::encode (morton_encoder_3d.hpp:35-37) or HilbertEncoder3D::encode
 * (hilbert_encoder_3d.hpp:29-66). keys_out (nullable): n keys in input order. */
int oct_encode(oct_ctx *ctx, int encoder, const double center[3], const double radii[3], uint64_t *keys_out);

/* Key algebra of single coordinates, batched on the device: PointEncoder::encode / decode
 * (morton_encoder_3d.hpp:35-46; hilbert_encoder_3d.hpp:29-66, 256-289). xyz: n x 3 u32 cell
 * coordinates (bits above 21 are ignored, as in the reference); keys: n u64. */
int oct_encode_coords(oct_ctx *ctx, int encoder, const uint32_t *xyz, size_t n, uint64_t *keys);
int oct_decode_keys(oct_ctx *ctx, int encoder, const uint64_t *keys, size_t n, uint32_t *xyz);

/* ---- H3 reorder: the radix passes of sortPoints (point_encoder.hpp:285-383) -- a stable sort by
 * key. perm_out[i] (nullable) = input index of the point now at sorted position i; keys_out
 * (nullable) = sorted keys. The device keeps the sorted cloud for the build and searches. */
int oct_sort(oct_ctx *ctx, uint32_t *perm_out, uint64_t *keys_out);
/* The reorder of sortPoints applied to caller records (point_encoder.hpp:321-323, 367-375: the points
 * and the PointMetadata vector move with the keys): out[i] = in[perm[i]] for the ctx's n records of
 * record_bytes bytes, perm the last oct_sort's. device_pointers != 0: in / out are device buffers
 * (gathered in place on the device); else host buffers (copied in, gathered on the device, copied
 * back). in and out must not overlap. */
int oct_gather_records(oct_ctx *ctx, const void *in, void *out, size_t record_bytes, int device_pointers);

/* One call for sortPoints(points, meta) (point_encoder.hpp:484-503): bbox + encode + sort, with the
 * same results as oct_bbox + oct_encode + oct_sort; the box stays on the device between the three
 * steps and is read back once at the end, so the host waits only there. */
int oct_sort_points(oct_ctx *ctx, int encoder, double center_out[3], double radii_out[3], uint32_t *perm_out,
                    uint64_t *keys_out);

/* Upload an already sorted cloud with its keys and Box (LinearOctree ctor inputs,
 * linear_octree.hpp:614-618), replacing what oct_sort left on the device. */
int oct_set_sorted(oct_ctx *ctx, const double *xyz, size_t n, size_t stride_bytes, const uint64_t *codes,
                   int encoder, const double center[3], const double radii[3]);
/* oct_set_sorted for a LinearOctree<PointsSoA> (linear_octree.hpp:614-618 over
 * point_containers.hpp:72-163): the sorted cloud as three host coordinate arrays. */
int oct_set_sorted_soa(oct_ctx *ctx, const double *x, const double *y, const double *z, size_t n, const uint64_t *codes,
                       int encoder, const double center[3], const double radii[3]);

/* ---- H4 build: LinearOctree(points, codes, box, enc) -> buildOctree (linear_octree.hpp:567-623);
 * maxPointsLeaf is the reference's global mainOptions.maxPointsLeaf (main_options.hpp:210). */
int oct_build(oct_ctx *ctx, uint32_t max_points_leaf, oct_tree_info *info);

/* Tree dump for parity checks, in the reference's node order (level-major, key-minor).
 * Any pointer may be NULL. prefixes: nTotal placeholder-bit keys (inter.prefixes); offsets: nTotal
 * first-child indices, 0 = leaf (offsets); ranges: nTotal x 2 [first, last) point ranges
 * (internalRanges); centers: nTotal x 3 (centers); radii: 22 x 3 (precomputedRadii). */
int oct_get_tree(oct_ctx *ctx, uint64_t *prefixes, uint32_t *offsets, uint64_t *ranges, double *centers,
                 double *radii);

/* ---- H5 radius searches: neighbors / neighborsPrune / neighborsStruct over the four kernels
 * (linear_octree.hpp:695-816, 1038-1087; kernel_factory.hpp:23-41).
 * Queries: either centres (centers != NULL, nq x 3 doubles) or indices into the sorted cloud
 * (center_idx != NULL), or, if both are NULL, all points 0..nq-1 of the sorted cloud.
 * radii: nq per-query radii, or NULL to use radius for every query (a double, as the kernels'
 * constructors take, kernel_3d.hpp:123).
 * counts_out (nq, nullable): |N(q)| (the harness keeps result.size(), neighbor_benchmarks.hpp:509).
 * Then oct_radius_fetch copies the result lists of the last search:
 *   neighbors/Prune -> indices (CSR, offsets nq+1), Struct -> [first,last) ranges (CSR of pairs),
 *   each list in the reference's own emission order. */
int oct_radius(oct_ctx *ctx, int kernel, int algo, int precision, const double *centers, const uint32_t *center_idx,
               const double *radii, double radius, size_t nq, int want_lists, uint32_t *counts_out,
               oct_search_stats *stats);
/* Allocation-free variant of oct_radius(want_lists = 1) + oct_radius_fetch, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference. It calls no hipMalloc / hipFree: the queries
 * (center_idx_dev, nq sorted-cloud indices on the device, or NULL for points 0 .. nq-1) and every
 * output are the caller's device buffers -- counts_dev (nq neighbour counts), offsets_dev (nq + 1
 * list offsets, in entries: indices, or [first, last) ranges for OCT_ALGO_STRUCT, u64), entries_dev
 * (u32 words: one per index, two per range) -- and the scan of the list sizes keeps its partial sums
 * in a fixed device array of the library (64 KiB per GPU). Limit: nq <= 67,108,864 per call. The
 * list total (indices or ranges) is returned in *total_out; an entries_capacity (u32 words) below
 * what the lists need fails with OCT_ERR_ARG after counts, offsets and *total_out are written, so
 * the caller can size its buffer and call again. Calls that overlap on one GPU (several contexts on
 * different streams) race on the shared scratch array. */
int oct_radius_lists_noalloc(oct_ctx *ctx, int kernel, int algo, int precision, const uint32_t *center_idx_dev,
                             double radius, size_t nq, uint32_t *counts_dev, uint64_t *offsets_dev, uint32_t *entries_dev,
                             size_t entries_capacity, uint64_t *total_out);
/* Vector-radii overloads (linear_octree.hpp:1044-1060, 1069-1075, 1082-1087; kernel_factory.hpp:33-41):
 * cube or square kernels with per-axis half extents, box = centre -/+ radii (kernel_abstract.hpp:39-45).
 * radii3: nq x 3 per-query half extents, or NULL to use radius3 for every query. Queries and lists
 * as for oct_radius (the lists are fetched with oct_radius_fetch). Comparisons only: no precision
 * argument. Other kernels return OCT_ERR_UNSUPPORTED, as the reference's static_assert rejects them. */
int oct_radius_box(oct_ctx *ctx, int kernel, int algo, const double *centers, const uint32_t *center_idx,
                   const double *radii3, const double radius3[3], size_t nq, int want_lists, uint32_t *counts_out,
                   oct_search_stats *stats);
/* After oct_radius(want_lists=1): offsets (nq+1) and entries (offsets[nq] u32 indices, or 2 u32 per range).
 * entries == NULL fetches the offsets only (size query). */
int oct_radius_fetch(oct_ctx *ctx, uint64_t *offsets, uint32_t *entries, size_t entries_capacity);
/* Device-resident variant for throughput runs: counts stay on the device (counts_dev may be NULL). */
int oct_radius_dev(oct_ctx *ctx, int kernel, int algo, int precision, const uint32_t *center_idx_dev, double radius,
                   size_t nq, uint32_t *counts_dev, oct_search_stats *stats);

/* neighborsApprox (linear_octree.hpp:981-1035): neighborsStruct with a depth cutoff. A node deeper
 * than `level` (the reference's uint32_t precisionLevel; getPrecisionLevel derives it from a
 * tolerance, and the (Point, radius, double tolerancePercentage, bool) overload passes the tolerance
 * itself as the level, linear_octree.hpp:1089-1093) is added whole when upper_bound != 0 and dropped
 * otherwise, without a kernel test. Queries, radii, precision and lists (Struct ranges, fetched with
 * oct_radius_fetch) as for oct_radius; radius3 non-NULL: cube / square with per-axis half extents
 * (then the comparisons run in FP64). */
int oct_radius_approx(oct_ctx *ctx, int kernel, int precision, const double *centers, const uint32_t *center_idx,
                      const double *radii, double radius, const double radius3[3], size_t nq, uint32_t level, int upper_bound,
                      int want_lists, uint32_t *counts_out, oct_search_stats *stats);

/* ---- H6 kNN: LinearOctree::knn (linear_octree.hpp:899-942), benchmark label KNNV2. Exact k
 * nearest under the (squared FP64 distance, sorted index) order; returns min(n, k) per query,
 * ascending; k <= 1024. idx_out: nq x k int32 (-1 pads), d2_out: nq x k FP64 squared distances
 * (NaN pads). OCT_PREC_FP32_RECHECK screens points with leaf-recentred FP32 distances and computes
 * the FP64 distance only for points that may beat the current k-th; results are identical. */
int oct_knn(oct_ctx *ctx, int precision, const double *centers, const uint32_t *center_idx, size_t nq, uint32_t k,
            int32_t *idx_out, double *d2_out, oct_search_stats *stats);
/* Device-resident kNN for throughput runs, same neighbour sets in the compact output form: int32
 * indices and FP32 squared distances (the exact FP64 distance rounded to nearest), nq x k each,
 * -1 / NaN past min(n, k). Queries are the sorted points first .. first + nq - 1
 * (center_idx_dev NULL), or nq device-resident indices into the sorted cloud (each < n; first = 0),
 * processed in the given order: ascending indices keep each wave's queries together. idx_dev /
 * d2_dev NULL: the results stay in the context and are read back with oct_knn_fetch (queries
 * [q0, q0 + nq) of the last such call; either output may be NULL). */
int oct_knn_dev(oct_ctx *ctx, int precision, const uint32_t *center_idx_dev, size_t first, size_t nq, uint32_t k,
                int32_t *idx_dev, float *d2_dev, oct_search_stats *stats);
int oct_knn_fetch(oct_ctx *ctx, size_t q0, size_t nq, int32_t *idx_out, float *d2_out);
/* Verification of kNN results that stay on the device (no reference counterpart: it compares result
 * sets too large to copy out, e.g. FP32 + recheck against FP64 for all points at k = 1000). Over nq
 * rows of k entries -- the caller's device buffers, or with both NULL the context's kept results of
 * the last oct_knn_dev -- out[0] = order-dependent 64-bit checksum of (position, index, FP32 d^2
 * bits), out[1] = rows whose distances are not ascending or whose indices lie outside the cloud,
 * out[2] = sum of the indices (mod 2^64). */
int oct_knn_checksum(oct_ctx *ctx, const int32_t *idx_dev, const float *d2_dev, size_t nq, uint32_t k, uint64_t out[3]);

/* ---- tuning (no counterpart in the reference: its per-query loop has no tile shape to choose).
 * OCT_TUNE_RADIUS_GROUP: queries per wave of the count-mode radius kernel (16, 32 or 64 queries
 * share one wave's traversal, 64 / G lanes serve each query), or 0 for the one-lane-per-query
 * kernel; -1 restores "autotuned, else 32". OCT_TUNE_RADIUS_MFMA: the point tests run on the vector
 * ALUs (a matrix-core form measured slower on every workload and is not built): 0 and -1 are
 * accepted, 1 fails with OCT_ERR_UNSUPPORTED.
 * oct_autotune_radius: times each candidate (G = 0, 16, 32, 64) on sample_nq queries taken from the
 * sorted cloud (16 equal strided chunks of consecutive SFC points; sample_nq >= n uses every point),
 * stores the fastest for (kernel, algo, precision), returns its G in *best and the milliseconds per G
 * in ms_out[0..3].
 * oct_tuned_radius: the G a count-mode search would use now (*mfma, if given, is always 0).
 * OCT_TUNE_RADIUS_REMAP: 1 (default, also -1) lets the count-mode vector point tests move the
 * queries of a leaf that at most half of a wave's queries want onto all 64 lanes; 0 keeps every
 * query on its own lanes.
 * OCT_TUNE_KNN_LANES: lanes per query of the kNN kernel (64 / lanes queries share one wave's
 * traversal); each k size class (k <= 16, 32, 64, ... 1024) offers up to three shapes, the
 * default first (k <= 32: 4, 2, 8; <= 64: 8, 4, 16; <= 128: 16, 8, 32; <= 256: 64, 32, 16;
 * above: 64); a value a class does not offer makes its kNN calls fail with OCT_ERR_ARG; -1
 * restores "autotuned, else the default". oct_autotune_knn times the offered shapes on sample_nq
 * all-points queries (16 strided chunks, as above; 0 = 2^18), stores the fastest for (precision,
 * size class of k), returns it in *best and the best milliseconds per shape in ms_out[0..2].
 * oct_tuned_knn: the lanes per query a kNN call with this k would use now.
 * OCT_TUNE_SORT: how oct_sort sorts more than 2048 (4096) keys: 1 orders them by their top 16 bits
 * in two device-wide passes, then sorts each 16-bit segment inside one workgroup (2048 keys per LDS
 * sort), which also writes the sorted cloud; 2 the same with 4096-key LDS sorts; 0 runs the
 * device-wide LSD pass per 8-bit digit and a separate gather. A 16-bit segment above 32768 keys
 * selects the LSD passes. -1 (the default) is 1, except that once a sort of the current points has
 * fallen back to the LSD passes, later sorts of the same points run them directly (new points reset
 * this). Results never depend on the tuning. */
enum oct_tune_knob {
    OCT_TUNE_RADIUS_GROUP = 1,
    OCT_TUNE_RADIUS_MFMA = 2,
    OCT_TUNE_KNN_LANES = 3,
    OCT_TUNE_RADIUS_REMAP = 4,
    OCT_TUNE_SORT = 5
};
int oct_set_tuning(oct_ctx *ctx, int knob, int value);
int oct_autotune_radius(oct_ctx *ctx, int kernel, int algo, int precision, double radius, size_t sample_nq, int *best,
                        double ms_out[4]);
int oct_tuned_radius(oct_ctx *ctx, int kernel, int algo, int precision, int *group, int *mfma);
int oct_autotune_knn(oct_ctx *ctx, int precision, uint32_t k, size_t sample_nq, int *best, double ms_out[3]);
int oct_tuned_knn(oct_ctx *ctx, int precision, uint32_t k, int *lanes);

/* ---- batches of independent clouds (no counterpart in the reference: its LinearOctree holds one
 * cloud and its harness runs clouds one after another). Each added cloud is encoded, sorted and built
 * exactly as by oct_sort_points + oct_build on its own context (same keys, permutation and tree);
 * oct_batch_build then lays all clouds out as one forest on the device, and each search call runs
 * every cloud's queries in one launch per search kind (kNN: one per k size class). Queries name a
 * cloud and a centre by its index in that cloud's sorted order; results use the same indices.
 * oct_batch_add copies the points (stride as oct_set_points), before any build and before
 * oct_batch_queries (OCT_ERR_STATE after them); encoder OCT_ENC_MORTON3D or
 * OCT_ENC_HILBERT3D; *cloud_id receives the cloud's number (0, 1, ...). oct_batch_radius: counts of
 * a count-mode search with a radius per query. oct_batch_knn: exact k nearest with a k per query
 * (1 .. kmax, kmax <= 1024); rows of kmax entries, -1 / NaN past min(n, k). oct_batch_info: a cloud's
 * size and tree after oct_batch_build. */
typedef struct oct_batch oct_batch;
int oct_batch_create(oct_batch **batch, int device);
int oct_batch_destroy(oct_batch *batch);
const char *oct_batch_last_error(const oct_batch *batch);
int oct_batch_add(oct_batch *batch, const double *xyz, size_t n, size_t stride_bytes, int encoder, uint32_t max_points_leaf,
                  int *cloud_id);
int oct_batch_build(oct_batch *batch);
/* Device-driven build of every cloud at once, with the same keys, permutations and trees as
 * oct_batch_build: per-cloud boxes and keys, one sort segmented by cloud, and a level-synchronous
 * build of the whole forest (each level of every cloud in one launch), enqueued with no host
 * synchronisation between the stages. With use_graph != 0 the sequence is captured in a HIP graph
 * on the first call and the graph is replayed on later calls; call again after oct_batch_set_points
 * to rebuild (the point counts stay). *ms (may be NULL) receives the device time of the sequence. */
int oct_batch_build_device(oct_batch *batch, int use_graph, double *ms);
/* Replaces the points of one cloud (same count; stride as oct_set_points) of a batch that is not
 * built yet or was built by oct_batch_build_device; the batch must be built again before searching. */
int oct_batch_set_points(oct_batch *batch, int cloud, const double *xyz, size_t stride_bytes);
/* Parity dumps of one cloud of a built batch: its sorted keys and permutation (indices into the
 * cloud's points as added), and (oct_batch_build_device only) its tree in the reference's node
 * order, with the arrays and NULL rules of oct_get_tree. */
int oct_batch_cloud_sorted(oct_batch *batch, int cloud, uint64_t *keys, uint32_t *perm);
int oct_batch_cloud_tree(oct_batch *batch, int cloud, uint64_t *prefixes, uint32_t *offsets, uint64_t *ranges,
                         double *centers, double *radii);
/* The search pass of a batch built on the device. oct_batch_queries copies a query set to the device
 * (radius queries: cloud, centre index in that cloud's sorted order, radius; kNN queries: cloud,
 * centre index, k in 1 .. kmax). oct_batch_pipeline then runs the whole sequence: the device build
 * of oct_batch_build_device, the grouping of the queries on the device (sorted by cloud and centre,
 * cut into wave tiles per cloud: 64 radius queries, 64 / lanes kNN queries of one k size class), the
 * counts of a count-mode radius search (kernel, algo NEIGHBORS / PRUNE / STRUCT, precision) and the
 * exact kNN search, each by one work-queue kernel whose resident waves take tiles from a device
 * counter. With use_graph != 0 the sequence is captured in a HIP graph once per shape (kernel,
 * algorithm, precision, lanes) and replayed; *ms (may be NULL) receives its device time.
 * oct_batch_pipeline_results copies the last pass's counts (nr) and kNN rows (nk x kmax, indices in
 * the query's cloud, -1 / NaN past min(n, k)); any pointer may be NULL.
 * oct_batch_autotune_knn times every lanes-per-query shape of each k size class present on the
 * batch's own kNN queries and keeps the fastest (lanes[c] for class c: k <= 16 << c; ms[3c + o]: the
 * time of option o, -1 when not run); the next pipeline run uses them. */
int oct_batch_queries(oct_batch *batch, const uint32_t *r_cloud, const uint32_t *r_idx, const double *r_radius, size_t nr,
                      const uint32_t *k_cloud, const uint32_t *k_idx, const uint32_t *k_k, size_t nk, uint32_t kmax);
int oct_batch_pipeline(oct_batch *batch, int kernel, int algo, int precision, int use_graph, double *ms);
int oct_batch_pipeline_results(oct_batch *batch, uint32_t *counts, int32_t *idx, double *d2);
int oct_batch_autotune_knn(oct_batch *batch, int precision, int lanes[7], double ms[21]);
int oct_batch_info(const oct_batch *batch, int cloud, size_t *n, oct_tree_info *info);
/* device milliseconds of the last batch searches, as oct_stage_times ([5] radius, [6] kNN: the last class) */
int oct_batch_stage_times(const oct_batch *batch, double ms[8]);
int oct_batch_radius(oct_batch *batch, int kernel, int algo, int precision, const uint32_t *cloud, const uint32_t *center_idx,
                     const double *radii, size_t nq, uint32_t *counts_out, oct_search_stats *stats);
int oct_batch_knn(oct_batch *batch, int precision, const uint32_t *cloud, const uint32_t *center_idx, const uint32_t *k,
                  size_t nq, uint32_t kmax, int32_t *idx_out, double *d2_out, oct_search_stats *stats);

/* ---- LAS ingest (LasFileReaderParallel::read / readMeta, readers/las_file_reader_parallel.hpp:23-145;
 * the reader handler, readers/handlers.hpp:57-72). Uncompressed LAS 1.0-1.4, point data formats 0-10,
 * extra bytes allowed; LAZ is rejected (OCT_ERR_UNSUPPORTED). oct_las_read_header parses the public
 * header block (n_points: the legacy count, or the LAS 1.4 64-bit count when the legacy one is 0).
 * oct_las_load makes the file's points the context's cloud (as oct_set_points, input order):
 * x = X * scale[0] + offset[0] in FP64 (:88-92), and decodes each point's attributes (:124-137,
 * PointMetadata, point_metadata.hpp:9-105) on the device. oct_las_fetch copies the loaded cloud
 * (n x 3) and / or the attributes (n records) to the host; either pointer may be NULL. */
typedef struct oct_las_header {
    uint8_t version_major, version_minor;
    uint8_t point_format; /* 0 .. 10 */
    uint8_t pad0;
    uint16_t record_length; /* bytes per point record (extra bytes included) */
    uint16_t header_size;
    uint32_t offset_to_points;
    uint32_t n_vlr;
    uint64_t n_points;
    double scale[3], offset[3];
    double min[3], max[3]; /* as stored in the header */
} oct_las_header;
typedef struct oct_point_meta {
    double intensity;
    uint8_t return_number, number_of_returns, scan_direction, edge_of_flight_line;
    uint8_t classification;
    int8_t scan_angle_rank;
    uint8_t user_data, pad0;
    uint16_t point_source_id, pad1;
    uint32_t r, g, b;
} oct_point_meta;
int oct_las_read_header(const char *path, oct_las_header *header);
int oct_las_load(oct_ctx *ctx, const char *path, oct_las_header *header);
int oct_las_fetch(oct_ctx *ctx, double *xyz, oct_point_meta *meta);

/* ---- timing: device milliseconds of the stages of the last calls (HIP events on the context's
 * stream): [0] bbox, [1] encode, [2] sort, [3] gather, [4] build, [5] search, [6] knn. */
int oct_stage_times(const oct_ctx *ctx, double ms[8]);
/* Measurement aid (no reference counterpart): the best of 5 timed device copies of `bytes` bytes
 * (16-byte words), in ms -- the practical read + write ceiling for the HBM-bound stages. */
int oct_copy_bandwidth(oct_ctx *ctx, size_t bytes, double *ms_out);

#ifdef __cplusplus
}
#endif
#endif
