// H4: linear-octree build on gfx950.
//
// The reference grows a cornerstone leaf array by rounds of split decisions (linear_octree.hpp:
// 151-333, 458-489), then derives the level-ordered node arrays (offsets, internalRanges, centers)
// by prefix sort, linking and recursion (:336-565). Starting from the root, its merges never fire,
// so its tree is the canonical one: a node splits iff it holds more than maxPointsLeaf points and
// its level is below 21, and nodes are numbered level by level, in key order within a level.
//
// This build produces those arrays directly, top-down, one level per step:
//   level l nodes -> (flag = count > M && l < 21) -> exclusive scan of flags -> children of the
//   flagged nodes are the next level, in key order, at indices levelStart[l+1] + 8 * rank + octant.
// A child's point range comes from two binary searches of the sorted keys inside its parent's
// range (the reference's lower_bound counting, :317-321); its FP64 centre is getCenter
// (point_encoder.hpp:517-539) evaluated in the reference's operation order.
// A last pass writes the leaf-recentred FP32 copy of the cloud used by the FP32 searches.
#include "tree_common.hpp"

namespace oct {

struct BuildGeom {
    double boxmin[3];
    double hl[3];
    double rad[22][3];
    int encoder;
};

__device__ inline void write_node(NodeRec *__restrict__ nodes, uint64_t *__restrict__ nkey, uint32_t u, uint64_t key,
                                  uint32_t level, uint32_t first, uint32_t last, const BuildGeom &g) {
    put_node(nodes, nkey, u, key, level, first, last, g.encoder, g.boxmin, g.hl, g.rad[level]);
}

__global__ void k_root(NodeRec *nodes, uint64_t *nkey, uint32_t *flags, uint32_t n, uint32_t max_leaf, BuildGeom g) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        write_node(nodes, nkey, 0, 0ull, 0, 0, n, g);
        flags[0] = (n > max_leaf) ? 1u : 0u;
    }
}

// Nodes of level `level` (indices [ls, le)), 8 lanes each. Flagged nodes write their 8 children.
__global__ __launch_bounds__(256) void k_children(const uint64_t *__restrict__ keys, NodeRec *__restrict__ nodes,
                                                 uint64_t *__restrict__ nkey, uint32_t *__restrict__ child,
                                                 const uint32_t *__restrict__ flags, const uint32_t *__restrict__ rank,
                                                 uint32_t *__restrict__ next_flags, uint32_t ls, uint32_t le, uint32_t level,
                                                 uint32_t max_leaf, BuildGeom g) {
    // 8 lanes per node (lane k: child k); a node's lanes leave or stay together
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i = t >> 3;
    const uint32_t k = (uint32_t)(t & 7u);
    if (i >= le - ls) return;
    const uint32_t u = ls + (uint32_t)i;
    if (!flags[i]) {
        if (k == 0) child[u] = 0;
        return;
    }
    const uint32_t cb = le + 8u * rank[i];
    if (k == 0) child[u] = cb;
    const NodeRec pr = nodes[u];
    const uint64_t pkey = nkey[u];
    const uint32_t cl = level + 1;
    const uint64_t step = node_range(cl);
    // child k ends where the keys reach child k + 1's prefix; it starts where child k - 1 ends (the
    // search over the whole parent range finds the same index as one starting at child k's start)
    const uint32_t hi = (k == 7) ? pr.last : lower_bound_keys(keys, pr.first, pr.last, pkey + (uint64_t)(k + 1) * step);
    const uint32_t prev = __shfl_up(hi, 1, 8);
    const uint32_t lo = (k == 0) ? pr.first : prev;
    write_node(nodes, nkey, cb + k, pkey + (uint64_t)k * step, cl, lo, hi, g);
    next_flags[cb - le + k] = (hi - lo > max_leaf && cl < kMaxDepth) ? 1u : 0u;
}

// Leaf-recentred FP32 coordinates, kRecentreLanes lanes per node (recentre_leaf); internal nodes'
// groups leave at once.
__global__ __launch_bounds__(256) void k_recentre(const double4 *__restrict__ p64, const NodeRec *__restrict__ nodes,
                                                 const uint32_t *__restrict__ child, uint32_t n_total, float4 *__restrict__ p32,
                                                 float *__restrict__ amax) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t u = t / kRecentreLanes;
    const uint32_t g = (uint32_t)(t % kRecentreLanes);
    if (u >= n_total) return;
    if (child[u] != 0) return;
    const float m = recentre_leaf<kRecentreLanes>(p64, nodes[u], g, p32);
    if (g == 0) amax[u] = m;
}

// Per internal node u: its children regrouped by corner bits, 8 lanes per node (node_ext_child).
struct LevelTable {
    uint32_t start[kMaxDepth + 3];
};
__global__ __launch_bounds__(256) void k_node_ext(const NodeRec *__restrict__ nodes, const uint64_t *__restrict__ nkey,
                                                 const uint32_t *__restrict__ child, const float *__restrict__ amax,
                                                 uint32_t n_total, int encoder, LevelTable lt, NodeExt *__restrict__ ext) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t u = t >> 3;
    if (u >= n_total) return;
    const uint32_t fc = child[u];
    if (fc == 0) return;
    uint32_t level = 0;
    while (level + 1 < kMaxDepth + 2 && lt.start[level + 1] <= u) level++;
    node_ext_child(nodes, nkey, child, amax, fc, (uint32_t)(t & 7u), level, encoder, ext + ((fc - 1) >> 3));
}

template <class T>
static hipError_t grow_preserve(oct_ctx *ctx, DevBuf<T> &b, size_t need, size_t valid) {
    if (need <= b.cap) return hipSuccess;
    size_t cap = b.cap ? b.cap : 1024;
    while (cap < need) cap *= 2;
    T *p = nullptr;
    hipError_t e = hipMalloc((void **)&p, cap * sizeof(T));
    if (e != hipSuccess) return e;
    if (b.p && valid) {
        e = hipMemcpyAsync(p, b.p, valid * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
    }
    if (b.p) (void)hipFree(b.p);
    b.p = p;
    b.cap = cap;
    return hipSuccess;
}

hipError_t launch_build(oct_ctx *ctx, uint32_t M, std::string &err) {
    const uint32_t n = (uint32_t)ctx->n;
    BuildGeom g;
    g.encoder = ctx->encoder;
    for (int a = 0; a < 3; a++) g.boxmin[a] = ctx->boxmin[a], g.hl[a] = ctx->half_lengths[a];
    for (int l = 0; l < 22; l++)
        for (int a = 0; a < 3; a++) g.rad[l][a] = ctx->level_radii[l][a];

    hipError_t e;
    // geometry for the searches: level radii, then per axis |centre| + radius (the kNN pruning margin);
    // copied from the context's own array, which outlives the copy
    for (int l = 0; l < 22; l++)
        for (int a = 0; a < 3; a++) ctx->geo[3 * l + a] = ctx->level_radii[l][a];
    for (int a = 0; a < 3; a++) ctx->geo[66 + a] = fabs(ctx->center[a]) + fabs(ctx->radii[a]);
    if ((e = ctx->radtab.reserve(69)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(ctx->radtab.p, ctx->geo, 69 * sizeof(double), hipMemcpyHostToDevice, ctx->stream)) !=
        hipSuccess)
        return e;
    size_t est = 64 + 4 * (size_t)n / (M ? M : 1);
    if ((e = grow_preserve(ctx, ctx->nodes, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->nkey, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->child, est + 1, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->flags, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->nrank, est, 0)) != hipSuccess) return e;
    if ((e = ctx->misc.reserve(12 * 1024 + 64)) != hipSuccess) return e;
    uint32_t *total_dev = ctx->misc.p + 8;

    stage_begin(ctx, ST_BUILD);
    k_root<<<1, 64, 0, ctx->stream>>>(ctx->nodes.p, ctx->nkey.p, ctx->flags.p, n, M, g);
    uint32_t ls = 0, le = 1, level = 0, n_internal = 0;
    ctx->level_start[0] = 0;
    ctx->level_start[1] = 1;
    for (;; level++) {
        const uint32_t cnt = le - ls;
        if ((e = exclusive_scan_u32(ctx, ctx->flags.p + ls, ctx->nrank.p + ls, cnt, total_dev)) != hipSuccess) return e;
        uint32_t nint = 0;
        if ((e = hipMemcpyAsync(&nint, total_dev, sizeof(nint), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
        if (nint > cnt || (nint && level >= kMaxDepth)) {
            err = "build: inconsistent split count";
            return hipErrorUnknown;
        }
        const size_t newle = (size_t)le + 8ull * nint;
        if (newle >= (1ull << 31)) {
            err = "build: node count exceeds 2^31";
            return hipErrorInvalidValue;
        }
        if ((e = grow_preserve(ctx, ctx->nodes, newle, le)) != hipSuccess) return e;
        if ((e = grow_preserve(ctx, ctx->nkey, newle, le)) != hipSuccess) return e;
        if ((e = grow_preserve(ctx, ctx->child, newle + 1, le)) != hipSuccess) return e;
        if ((e = grow_preserve(ctx, ctx->flags, newle, le)) != hipSuccess) return e;
        if ((e = grow_preserve(ctx, ctx->nrank, newle, le)) != hipSuccess) return e;
        unsigned nb = (unsigned)(((size_t)cnt * 8 + 255) / 256);
        k_children<<<nb, 256, 0, ctx->stream>>>(ctx->sorted_keys, ctx->nodes.p, ctx->nkey.p, ctx->child.p, ctx->flags.p + ls,
                                                ctx->nrank.p + ls, ctx->flags.p + le, ls, le, level, M, g);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        n_internal += nint;
        ctx->level_start[level + 2] = (uint32_t)newle;
        if (nint == 0) break;
        ls = le;
        le = (uint32_t)newle;
    }
    const uint32_t n_total = le;
    // sentinel: offsets has nTotal + 1 entries in the reference (linear_octree.hpp:497)
    if ((e = hipMemsetAsync(ctx->child.p + n_total, 0, sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
    for (uint32_t l = level + 2; l < kMaxDepth + 3; l++) ctx->level_start[l] = n_total;
    ctx->n_total = n_total;
    ctx->n_internal = n_internal;
    ctx->n_leaf = n_total - n_internal;
    ctx->n_levels = level + 1;
    ctx->max_depth_seen = level;
    ctx->max_points_leaf = M;

    // leaf-recentred FP32 cloud
    if ((e = ctx->p32.reserve(ctx->n)) != hipSuccess) return e;
    if ((e = ctx->amax.reserve(n_total)) != hipSuccess) return e;

    if (n_total) {
        size_t threads = (size_t)n_total * kRecentreLanes;
        unsigned nb = (unsigned)((threads + 255) / 256);
        k_recentre<<<nb, 256, 0, ctx->stream>>>(ctx->p64.p, ctx->nodes.p, ctx->child.p, n_total, ctx->p32.p, ctx->amax.p);
    }
    // children regrouped by corner bits, for the count-mode searches
    if ((e = ctx->next.reserve(n_internal ? n_internal : 1)) != hipSuccess) return e;
    if (n_internal) {
        LevelTable lt;
        for (uint32_t l = 0; l < kMaxDepth + 3; l++) lt.start[l] = ctx->level_start[l];
        unsigned nb = (unsigned)(((size_t)n_total * 8 + 255) / 256);
        k_node_ext<<<nb, 256, 0, ctx->stream>>>(ctx->nodes.p, ctx->nkey.p, ctx->child.p, ctx->amax.p, n_total, ctx->encoder, lt,
                                                ctx->next.p);
    }
    stage_end(ctx, ST_BUILD);
    return hipGetLastError();
}

}  // namespace oct
