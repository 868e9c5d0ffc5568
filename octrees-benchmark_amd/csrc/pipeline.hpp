// Device-driven batch build: data layout shared by pipeline.hip (kernels) and batch.hip (host).
#pragma once

#include "internal.hpp"

namespace oct {

// Per cloud, on the device: its Box (mbb: centre, radii), octree geometry and build parameters.
struct CloudGeom {
    double center[3], radii[3];
    double boxmin[3], hl[3];
    double rad[22][3];
    uint32_t encoder, max_leaf;
    uint32_t pad[2];
};

// A bbox / encode work unit: [start, start + len) inside cloud `cloud`.
struct Chunk {
    uint32_t start, len, cloud, pad;
};
constexpr uint32_t kChunkPoints = 1u << 16;

constexpr uint32_t kNoCloud = 0xffffffffu;  // padding nodes of level 0
constexpr uint32_t kLvCount = kMaxDepth + 3;  // level starts 0 .. 23; [23] = node count
// control words: sort tile counters, sort look-back error, build error bits, level starts
constexpr uint32_t kCtrlSortCounters = 0, kCtrlSortErr = 8, kCtrlBuildErr = 9, kCtrlLv = 16;
constexpr uint32_t kCtrlWords = 64;
constexpr uint32_t kBuildErrDepth = 1u, kBuildErrCapacity = 2u;

// Device pointers and sizes of one batch build. Sizes fixed at enqueue time are the host's
// (points, clouds, chunks, tiles, node capacity); everything the build discovers stays on the device.
struct ForestBuild {
    size_t n = 0;            // points of all clouds
    uint32_t C = 0;          // clouds
    uint32_t R = 0;          // level-0 slots (C padded to 1 mod 8)
    uint32_t cap = 0;        // node capacity (child has cap + 1 entries, ext cap / 8 + 1)
    const double *xyz = nullptr;  // n x 3, clouds concatenated
    const Chunk *chunks = nullptr;
    uint32_t nchunks = 0;
    const uint32_t *cloud_chunk0 = nullptr;  // C + 1
    const uint32_t *cloud_off = nullptr;     // C + 1
    const uint32_t *cloud_enc = nullptr;     // C
    const uint32_t *cloud_leaf = nullptr;    // C
    const SortTile *tiles = nullptr;
    uint32_t ntiles = 0;
    CloudGeom *geom = nullptr;               // C
    double *radtab = nullptr;                // C x 66
    double *magtab = nullptr;                // C x 3
    uint32_t *roots = nullptr;               // C
    double *partial = nullptr;               // nchunks x 6
    uint64_t *keys[2] = {nullptr, nullptr};  // n each; sorted keys end in keys[0]
    uint32_t *vals[2] = {nullptr, nullptr};  // n each; the permutation (global positions) ends in vals[0]
    uint32_t *hist = nullptr;                // C x kSegBaseStride
    uint32_t *bases = nullptr;               // C x kSegBaseStride
    uint32_t *lookback = nullptr;            // kRadixPassCount x ntiles x 256
    uint32_t *ctrl = nullptr;                // kCtrlWords
    uint32_t *part = nullptr;                // level block sums (512)
    uint32_t *cl_internal = nullptr;         // C
    uint32_t *cl_depth = nullptr;            // C
    double4 *p64 = nullptr;                  // n
    float4 *p32 = nullptr;                   // n
    NodeRec *nodes = nullptr;                // cap
    uint64_t *nkey = nullptr;                // cap
    uint32_t *child = nullptr;               // cap + 1
    uint32_t *flags = nullptr;               // cap
    uint32_t *ncloud = nullptr;              // cap
    float *amax = nullptr;                   // cap
    NodeExt *ext = nullptr;                  // cap / 8 + 1
};

hipError_t enqueue_forest_build(hipStream_t s, const ForestBuild &F);

// ---- queries of a batch search pass, grouped on the device
constexpr int kKnnClasses = 7;  // kNN size classes: k <= 16, 32, ..., 1024
// control words of one query kind: sort tile counters [0, 8), sort error [8], per-class wave-tile
// counts [16, 24), per-class work-queue counters [32, 40), a zero word [48]
constexpr uint32_t kQCtrlTiles = 16, kQCtrlQueue = 32, kQCtrlZero = 48;

// One query kind: radius queries (one class, kq null) or kNN queries (classes by k). Every run sorts
// the queries by (class, cloud, centre index), finds each (class, cloud) group's range, cuts the
// groups into wave tiles of G[class] slots and writes, per class, the slot -> query table (order)
// and the tile -> (cloud | valid slots << 20) table that the search kernels' work queues read.
struct QueryGroups {
    uint32_t n = 0;       // queries
    uint32_t nclass = 1;  // 1 (radius) or kKnnClasses
    const uint32_t *cloud = nullptr, *idx = nullptr, *kq = nullptr;
    uint64_t *keys[2] = {nullptr, nullptr};  // n each
    uint32_t *vals[2] = {nullptr, nullptr};  // n each
    uint32_t *gidx = nullptr;      // per query: its centre's index in the concatenated sorted clouds
    uint32_t *hist = nullptr;      // kRadixPassCount x 256
    uint32_t *bases = nullptr;     // kRadixPassCount x 256
    uint32_t *lookback = nullptr;  // 5 x tiles x 256
    uint32_t *ctrl = nullptr;      // kCtrlWords
    uint32_t *gstart = nullptr;    // nclass x C + 1
    uint32_t *toff = nullptr;      // nclass x C
    uint32_t G[kKnnClasses] = {64, 64, 64, 64, 64, 64, 64};  // slots per wave tile
    uint32_t slots[kKnnClasses] = {0};                      // slot capacity per class
    uint32_t *order[kKnnClasses] = {nullptr};
    uint32_t *tile_info[kKnnClasses] = {nullptr};
};
// the slot capacity a class needs whatever the queries: (ceil(n / G) + C) tiles of G slots
inline uint32_t group_slots(uint32_t n, uint32_t C, uint32_t G) { return ((n + G - 1) / G + C) * G; }
hipError_t enqueue_query_groups(hipStream_t s, const QueryGroups &Q, uint32_t C, const uint32_t *cloud_off);
// k -> size class (16 << class >= k)
__host__ __device__ inline uint32_t knn_class_of(uint32_t k) {
    uint32_t c = 0;
    while (c + 1 < (uint32_t)kKnnClasses && (16u << c) < k) c++;
    return c;
}

}  // namespace oct
