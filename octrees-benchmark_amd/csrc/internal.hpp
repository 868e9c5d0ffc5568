// Internal declarations shared by the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/octree_amd.h"

namespace oct {

constexpr uint32_t kMaxDepth = 21;                       // morton_encoder_3d.hpp:23
constexpr uint64_t kUpperBound = 0x8000000000000000ull;  // morton_encoder_3d.hpp:29
constexpr uint32_t kMaxCoord = (1u << kMaxDepth) - 1u;

// Node record read by the searches: FP64 centre (computeGeometry, linear_octree.hpp:557-565) and
// the [first, last) point range (internalRanges). 32 bytes, so one aligned scalar load fetches it.
struct alignas(32) NodeRec {
    double cx, cy, cz;
    uint32_t first, last;
};

// Children of one internal node, in "combo" order c = bx | by << 1 | bz << 2, where (bx, by, bz) are
// the child's corner bits at its level. The 8 children take only two centre values per axis, so
// ax holds them: child c's centre is (ax[bx], ax[2 + by], ax[4 + bz]), bit-identical to its own
// FP64 centre (same corner integers, same formula). Indexed by (first child - 1) / 8 of the node.
struct alignas(64) NodeExt {
    double ax[6];       // x lo / hi, y lo / hi, z lo / hi centre coordinates of the children
    float amax[8];      // combo c: max |a|^2 of the leaf's recentred FP32 points (leaves only)
    uint32_t first[8];  // combo c: point range [first, last)
    uint32_t last[8];
    uint32_t gc[8];     // combo c: its first child, 0 = leaf
    uint32_t pad[4];
};
static_assert(sizeof(NodeExt) == 192, "NodeExt is three 64-byte lines");

// Device buffer that only grows.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t bytes = (n ? n : 1) * sizeof(T);
        hipError_t e = hipMalloc((void **)&p, bytes);
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Radix sort geometry: tiles of kSortTile keys, kRadixPassCount passes of 8 bits.
constexpr int kSortTile = 6144;
constexpr int kRadixPassCount = 8;
// A sort tile of a batch: [start, start + len) inside cloud `cloud`, whose first tile is `first`.
struct SortTile {
    uint32_t start, len, cloud, first;
};
// Per-cloud bin bases of a segmented sort: bases[cloud * kSegBaseStride + pass * 256 + digit].
constexpr uint32_t kSegBaseStride = kRadixPassCount * 256;
// Partition sort of one cloud (keys below 2^63): device-wide stable passes on bits 47..54 and 55..62,
// then workgroup sorts of the 16-bit segments, splitting big ones from bits 39..46 down. The encode
// kernel writes the histograms of bits 55..62 and 47..54 after the 8 LSD digit histograms.
constexpr uint32_t kMsdShift1 = 55, kMsdShift2 = 47, kMsdShift3 = 39;
constexpr int kEncHists = kRadixPassCount + 2;

// Bits of the context's device error word (oct_ctx::derr): kernels set them with atomicOr, the host
// reads and clears the word at its next synchronisation and fails the call.
constexpr uint32_t kErrStack = 1u;  // a traversal stack was full: a pending node would have been dropped

enum Stage { ST_BBOX = 0, ST_ENCODE, ST_SORT, ST_GATHER, ST_BUILD, ST_SEARCH, ST_KNN, ST_N };

}  // namespace oct

struct oct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // input cloud
    size_t n = 0;
    oct::DevBuf<double> xyz;  // n x 3, input order
    bool has_points = false;
    // LAS ingest: raw point records and the decoded per-point attributes of the last file loaded
    oct::DevBuf<unsigned char> las_rec;
    oct::DevBuf<oct_point_meta> las_meta;
    oct::DevBuf<unsigned char> rec_in, rec_out;  // oct_gather_records staging of host records
    size_t las_n = 0;
    bool las_loaded = false;

    // encode + sort
    int encoder = 0;
    double center[3] = {0, 0, 0}, radii[3] = {0, 0, 0};
    oct::DevBuf<uint64_t> keys[2];
    oct::DevBuf<uint32_t> vals[2];
    oct::DevBuf<uint32_t> hist;      // 8 x 256 digit histograms, then bits 55..62 and 47..54
    oct::DevBuf<double> box;         // oct_sort_points: min, max, centre and radii per axis, on the device
    oct::DevBuf<uint32_t> lookback;  // per pass, per tile, per bin
    oct::DevBuf<uint32_t> misc;      // tile counters, flags, scan partials
    oct::DevBuf<uint32_t> msd;             // partition sort: bases, prefix counts, segment jobs, counters
    int sort_override = -1;                // OCT_TUNE_SORT
    bool sort_lsd_learned = false;         // a default-form sort of these points fell back to the LSD passes
    bool p64_sorted = false;               // the last sort also wrote the sorted cloud (p64)
    bool hist_lsd = false;                 // the encode kernel wrote the 8 LSD digit histograms too
    oct::DevBuf<uint64_t> keys_p;          // partition sort: keys ordered by their top 16 bits
    oct::DevBuf<uint32_t> vals_p;
    uint64_t *sorted_keys = nullptr; // points into keys[0|1]
    uint32_t *perm = nullptr;        // points into vals[0|1]
    bool has_keys = false, has_sorted = false;

    // sorted cloud
    oct::DevBuf<double4> p64;  // sorted xyz, padded to 32 B
    oct::DevBuf<float4> p32;   // leaf-recentred FP32 xyz + |p|^2

    // tree (BFS / level order, as the reference's sorted node arrays)
    bool has_tree = false;
    uint32_t max_points_leaf = 0;
    uint32_t n_leaf = 0, n_internal = 0, n_total = 0, max_depth_seen = 0, n_levels = 0;
    uint32_t level_start[oct::kMaxDepth + 3] = {0};
    double boxmin[3] = {0, 0, 0};
    double half_lengths[3] = {0, 0, 0};
    double level_radii[22][3] = {{0}};
    double geo[69] = {0};           // level_radii (66) then |centre| + radius per axis (3)
    oct::DevBuf<double> radtab;     // geo on the device
    oct::DevBuf<oct::NodeRec> nodes;
    oct::DevBuf<uint32_t> child;    // first child index, 0 = leaf
    oct::DevBuf<uint64_t> nkey;     // node start key (SFC format)
    oct::DevBuf<uint32_t> flags;    // per-node split flags (build scratch)
    oct::DevBuf<uint32_t> nrank;    // per-node rank among split nodes of its level (build scratch)
    oct::DevBuf<float> amax;        // per-leaf max squared norm of its recentred FP32 points
    oct::DevBuf<oct::NodeExt> next;  // per internal node: its children in combo order
    oct::DevBuf<uint32_t> scan;     // scan scratch

    // searches
    oct::DevBuf<uint32_t> qidx;     // query order / indices
    oct::DevBuf<double> qxyz;       // explicit query centres
    oct::DevBuf<double> qrad;       // per-query radii
    oct::DevBuf<double> qrad3;      // per-query per-axis half extents (box kernels)
    oct::DevBuf<uint32_t> kc_xyz;   // key codec staging
    oct::DevBuf<uint64_t> kc_keys;
    oct::DevBuf<uint32_t> counts;   // per-query counts
    oct::DevBuf<uint64_t> qoff;     // per-query list offsets
    oct::DevBuf<uint32_t> entries;  // result lists
    oct::DevBuf<unsigned long long> stats;  // search counters
    size_t last_nq = 0;
    int last_algo = 0;
    int group_override = -1;        // OCT_TUNE_RADIUS_GROUP (-1: use tuned / default)
    int remap_override = -1;        // OCT_TUNE_RADIUS_REMAP (-1: default, on)
    int group_tuned[4][4][2];       // autotuned group per (kernel, algo, precision), -1 = not tuned
    bool has_lists = false;

    // kNN
    oct::DevBuf<int32_t> kidx;
    oct::DevBuf<double> kd2;
    oct::DevBuf<float> kd2f;        // oct_knn_dev results kept in the context
    size_t knn_nq = 0;              // rows of the kept oct_knn_dev results
    uint32_t knn_k = 0;
    int knn_lanes_override = -1;    // OCT_TUNE_KNN_LANES (-1: tuned / default)
    int knn_lanes_tuned[2][8];      // autotuned lanes per query per (precision, size class), -1 = not tuned

    oct::DevBuf<uint32_t> derr;     // device error word (kErr* bits), zero when no kernel failed

    hipEvent_t ev0[oct::ST_N] = {}, ev1[oct::ST_N] = {};
    bool stage_events_off = false;  // launch sequences captured in graphs record no stage events
    bool ev_used[oct::ST_N] = {};
};

namespace oct {

// ---- launchers implemented in the .hip files (all enqueue on ctx->stream) ----
hipError_t launch_bbox(oct_ctx *ctx, double mn[3], double mx[3]);
// The same reduction left on the device: ctx->box receives min, max, centre and radii (12 doubles),
// nothing is read back (n >= 1).
hipError_t launch_bbox_dev(oct_ctx *ctx);
// Encode with the box given on the host, or (center == nullptr) the one launch_bbox_dev left in ctx->box.
hipError_t launch_encode(oct_ctx *ctx, int encoder, const double center[3], const double radii[3]);
hipError_t launch_sort(oct_ctx *ctx, std::string &err);
hipError_t launch_gather(oct_ctx *ctx);
hipError_t launch_build(oct_ctx *ctx, uint32_t max_points_leaf, std::string &err);
// One radius-search launch. mode 0: per-query neighbour counts; mode 1: counts + list-entry counts
// (ecounts); mode 2: fill the lists at offsets. Query slot s (0..nq-1) processes query id order[s] (or s); the
// centre is qxyz[id] if given, else the sorted point qidx[id] (or id); outputs are indexed by id.
struct RadiusLaunch {
    int kernel = 1, algo = 3, precision = 0, mode = 0;
    int group = 16;  // count mode: queries per wave (16, 32, 64), or 0 for the one-lane-per-query kernel
    bool remap = true;  // count mode, vector point tests: compact the queries of sparsely wanted leaves
    uint32_t approx_level = ~0u;  // OCT_ALGO_APPROX: neighborsApprox depth cutoff and bound mode
    bool approx_upper = false;
    // count mode over a forest of clouds (batches): wave tiles of 64 slots, tile_info[tile] = cloud |
    // valid slots << 20, roots[cloud], radtab[cloud * 66 + 3 * level + axis]; null: the context's cloud
    const uint32_t *tile_info = nullptr;
    const uint32_t *roots = nullptr;
    const double *radtab = nullptr;
    // work-queue form (count mode over a forest, 64-slot tiles, vector ALUs, no stats): device-resident
    // tile count and a zeroed tile counter; the grid is the resident waves, not the tiles
    const uint32_t *queue_tiles = nullptr;
    uint32_t *queue_next = nullptr;
    const uint32_t *order = nullptr;
    const uint32_t *qidx = nullptr;
    const double *qxyz = nullptr;
    const double *qrad = nullptr;
    double radius = 0.0;
    const double *qrad3 = nullptr;  // box kernels with per-axis half extents: nq x 3, or
    double radius3[3] = {0, 0, 0};  // one triple for every query when use_r3 && !qrad3
    bool use_r3 = false;
    size_t nq = 0;
    uint32_t *counts = nullptr;
    uint32_t *ecounts = nullptr;
    uint64_t *ecounts64 = nullptr;
    const uint64_t *offsets = nullptr;
    uint32_t *entries = nullptr;
    unsigned long long *stats = nullptr;
};
hipError_t launch_radius(oct_ctx *ctx, const RadiusLaunch &L);
hipError_t launch_key_codec(oct_ctx *ctx, int encoder, bool decode, uint32_t *xyz_dev, uint64_t *keys_dev, size_t n);
// One kNN launch. Query slot s (0..nq-1) processes query id order[s] (or s); the centre is qxyz[id]
// if given, else the sorted point qidx[id] (or qoff + id); outputs (nq x k, row id) go to idx and
// to d2 (FP64) and / or d2f (FP32), either of which may be null.
struct KnnLaunch {
    int precision = 0;
    const uint32_t *order = nullptr;
    const uint32_t *qidx = nullptr;
    const double *qxyz = nullptr;
    uint32_t qoff = 0;
    size_t nq = 0;
    uint32_t k = 0;
    int32_t *idx = nullptr;
    double *d2 = nullptr;
    float *d2f = nullptr;
    unsigned long long *stats = nullptr;
    int lanes = 0;  // lanes per query (0: the size class's default)
    // forests of clouds (batches): per wave tile cloud | valid slots << 20, roots, per-cloud geometry
    // (66 level radii each, 3 magnitudes each) and per-query k; null: the context's one cloud
    const uint32_t *tile_info = nullptr;
    const uint32_t *roots = nullptr;
    const double *radtab = nullptr;
    const double *magtab = nullptr;
    const uint32_t *kq = nullptr;
    uint32_t kclass = 0;  // the largest k of the launch when kq is set (0: k)
    // forests: per-cloud start in the concatenated sorted clouds; set, indices are written local to
    // each query's cloud (null: forest-wide indices)
    const uint32_t *cloud_off = nullptr;
    // work-queue form: device-resident wave-tile count and a zeroed tile counter
    const uint32_t *queue_tiles = nullptr;
    uint32_t *queue_next = nullptr;
};
hipError_t launch_knn(oct_ctx *ctx, const KnnLaunch &L, std::string &err);
// lanes-per-query shapes offered for k (default first); returns how many
int knn_lane_options(uint32_t k, int out[3]);
// Batches: stable sort of every cloud's (key, value) range at once (8 passes; keys0 / vals0 in and
// out, the values of the first pass are the global positions); lookback (8 x ntiles x 256) and
// counters (8) must be zero.
hipError_t launch_sort_segmented(hipStream_t s, uint64_t *keys0, uint64_t *keys1, uint32_t *vals0, uint32_t *vals1,
                                 size_t n, const SortTile *tiles, size_t ntiles, const uint32_t *bases,
                                 uint32_t *lookback, uint32_t *counters, uint32_t *err);
// Stable sort of n (key, value) pairs by the low 8 * npasses key bits; values of the first pass are the
// positions; result in keys0 / vals0 for an even npasses, keys1 / vals1 for an odd one. bases: the
// exclusive bin offsets of every pass (pass * 256 + digit); lookback (npasses x tiles x 256) and
// counters (npasses) must be zero.
hipError_t launch_sort_passes(hipStream_t s, uint64_t *keys0, uint64_t *keys1, uint32_t *vals0, uint32_t *vals1, size_t n,
                              int npasses, const uint32_t *bases, uint32_t *lookback, uint32_t *counters, uint32_t *err);
// out[i] = in[perm[i]] for records of rb bytes (device buffers)
hipError_t launch_gather_records(hipStream_t s, const void *in, const uint32_t *perm, void *out, size_t n, size_t rb);
// out[i] = (xyz[3 perm[i]], xyz[3 perm[i] + 1], xyz[3 perm[i] + 2], 0)
hipError_t launch_gather_raw(hipStream_t s, const double *xyz, const uint32_t *perm, double4 *out, size_t n);
// whether oct_sort will take the partition sort for the context's current cloud and tuning (the encode
// kernel then writes only the partition histogram)
bool sort_partition_planned(const oct_ctx *ctx);
hipError_t exclusive_scan_u32_to_u64(oct_ctx *ctx, const uint32_t *in, uint64_t *out, size_t n);
hipError_t exclusive_scan_u32(oct_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev);

void stage_begin(oct_ctx *ctx, Stage s);
void stage_end(oct_ctx *ctx, Stage s);

}  // namespace oct
