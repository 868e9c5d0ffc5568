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
#include <algorithm>

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

// ---- device-driven level loop: no host round trip between levels
//
// Each level is two launches that take the level's node range [ls, le) from ctl[l] (written by the
// level above): k_level_scan ranks the split flags of 2048-node tiles with a chained scan (dynamic
// tile ids, decoupled look-back in which one wave reads 64 predecessor tiles per step, bounded spins),
// points every node at its first child (0 = leaf) and, in its last tile, writes ctl[l + 1];
// k_level_children then builds the children of the split nodes, 8 lanes per node over a grid-stride
// loop that fills the device (as k_children). Levels past the last one find an empty range and
// return, so the host enqueues a batch of levels and reads the level table back once. The node
// arrays are sized up front; a level that would outgrow them raises an error bit, and the host then
// runs the level loop with growing arrays instead.
constexpr uint32_t kLvlPer = 8, kLvlTile = 256 * kLvlPer;
constexpr uint32_t kLvlAgg = 1u << 30, kLvlInc = 2u << 30, kLvlMask = (1u << 30) - 1u, kLvlSpin = 1u << 24;
constexpr uint32_t kLvlErrSpin = 1u, kLvlErrCap = 2u;
constexpr int kLvlCount = kMaxDepth + 1;  // levels 0 .. 21 (level 21 nodes never split)
constexpr int kLvlFirstBatch = 13;        // levels enqueued before the first read-back (depth <= 12)

struct LevelCtl {
    uint32_t ls, le;
};

__device__ inline uint32_t lvl_wave_incl_scan(uint32_t v) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(v, off);
        if (l >= (uint32_t)off) v += o;
    }
    return v;
}

// root, level table (level 0 = the root, every other level empty until its parent level fills it),
// zero look-back words, tile counters and error word
__global__ __launch_bounds__(256) void k_level_init(NodeRec *nodes, uint64_t *nkey, uint32_t *flags, uint32_t n,
                                                    uint32_t max_leaf, BuildGeom g, LevelCtl *ctl,
                                                    uint32_t *__restrict__ lookback, size_t nlook,
                                                    uint32_t *__restrict__ counters, uint32_t *__restrict__ err,
                                                    uint32_t *__restrict__ total) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = t; i < nlook; i += stride) lookback[i] = 0u;
    if (t == 0) *total = 0u;
    if (t < (size_t)kLvlCount) counters[t] = 0u;
    if (t <= (size_t)kLvlCount) ctl[t] = t == 0 ? LevelCtl{0u, 1u} : LevelCtl{0u, 0u};
    if (t == 0) {
        write_node(nodes, nkey, 0, 0ull, 0, 0, n, g);
        flags[0] = (n > max_leaf) ? 1u : 0u;
        *err = 0u;
    }
}

__global__ __launch_bounds__(256) void k_level_scan(const uint32_t *__restrict__ flags, uint32_t *__restrict__ child,
                                                    LevelCtl *__restrict__ ctl, uint32_t level, uint32_t *__restrict__ lookback,
                                                    uint32_t *__restrict__ tile_counter, uint32_t cap, uint32_t *__restrict__ err,
                                                    uint32_t *__restrict__ total) {
    __shared__ uint32_t s_wsum[4], s_tile, s_prefix;
    const LevelCtl L = ctl[level];
    if (L.le <= L.ls) return;
    const uint32_t ntiles = (L.le - L.ls + kLvlTile - 1) / kLvlTile;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles) break;
        const uint32_t b0 = L.ls + tile * kLvlTile + tid * kLvlPer;  // this thread's nodes b0 .. b0 + 7
        uint32_t f[kLvlPer], sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < kLvlPer; i++) {
            f[i] = b0 + i < L.le ? flags[b0 + i] : 0u;
            sum += f[i];
        }
        const uint32_t incl = lvl_wave_incl_scan(sum);
        if (lane == 63) s_wsum[w] = incl;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (uint32_t ww = 0; ww < 4; ww++) {
            const uint32_t x = s_wsum[ww];
            pre += ww < w ? x : 0u;
            tot += x;
        }
        if (w == 0) {  // chained scan over the tiles: wave 0 looks back 64 tiles per step
            uint32_t excl = 0;
            if (tile == 0) {
                if (lane == 0) store_flag(&lookback[0], kLvlInc | tot);
            } else {
                if (lane == 0) store_flag(&lookback[tile], kLvlAgg | tot);
                int64_t base = (int64_t)tile - 1;  // lane i reads tile base - i (before tile 0: an empty prefix)
                uint32_t spins = 0;
                for (;;) {
                    const int64_t t = base - (int64_t)lane;
                    const uint32_t x = t >= 0 ? load_flag(&lookback[t]) : kLvlInc;
                    const uint64_t inc = ballot64((x >> 30) == 2u), wait = ballot64((x >> 30) == 0u);
                    const uint32_t first = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;  // nearest inclusive prefix
                    const uint64_t need = first >= 63u ? ~0ull : ((2ull << first) - 1ull);
                    if (wait & need) {  // a tile up to it has published nothing yet
                        if (++spins > kLvlSpin) {
                            if (lane == 0) atomicOr(err, kLvlErrSpin);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint32_t v = lane <= first ? (x & kLvlMask) : 0u;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                    excl += v;
                    if (first < 64u) break;
                    base -= 64;
                }
                if (lane == 0) store_flag(&lookback[tile], kLvlInc | ((excl + tot) & kLvlMask));
            }
            if (lane == 0) {
                s_prefix = excl;
                if (tile == ntiles - 1) {  // the level's split count: the next level's range
                    const uint64_t nle = (uint64_t)L.le + 8ull * (excl + tot);
                    if (nle <= cap) {
                        ctl[level + 1] = LevelCtl{L.le, (uint32_t)nle};
                    } else {
                        ctl[level + 1] = LevelCtl{L.le, L.le};
                        atomicOr(err, kLvlErrCap);
                    }
                    if (excl + tot == 0) *total = L.le;  // no node splits: the tree is complete
                }
            }
        }
        __syncthreads();
        uint32_t r = s_prefix + pre + incl - sum;  // rank of this thread's first split node in the level
#pragma unroll
        for (uint32_t i = 0; i < kLvlPer; i++) {
            if (b0 + i < L.le) child[b0 + i] = f[i] ? L.le + 8u * r : 0u;
            r += f[i];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_level_children(const uint64_t *__restrict__ keys, NodeRec *__restrict__ nodes,
                                                        uint64_t *__restrict__ nkey, const uint32_t *__restrict__ child,
                                                        uint32_t *__restrict__ flags, const LevelCtl *__restrict__ ctl,
                                                        uint32_t level, uint32_t cap, uint32_t max_leaf, BuildGeom g,
                                                        uint32_t *__restrict__ err) {
    const LevelCtl L = ctl[level];
    if (L.le <= L.ls || ctl[level + 1].le <= ctl[level + 1].ls) return;  // no node of this level splits
    const size_t lanes = (size_t)(L.le - L.ls) * 8, stride = (size_t)gridDim.x * blockDim.x;  // stride: a multiple of 8
    const uint32_t k = threadIdx.x & 7u, cl = level + 1;
    const uint64_t step = node_range(cl);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < lanes; t += stride) {
        const uint32_t u = L.ls + (uint32_t)(t >> 3);  // a node's 8 lanes move together
        const uint32_t cb = child[u];
        if (cb == 0) continue;
        if ((uint64_t)cb + 8u > cap) {
            if (k == 0) atomicOr(err, kLvlErrCap);
            continue;
        }
        const NodeRec pr = nodes[u];
        const uint64_t pkey = nkey[u];
        const uint32_t hi = (k == 7) ? pr.last : lower_bound_keys(keys, pr.first, pr.last, pkey + (uint64_t)(k + 1) * step);
        const uint32_t prev = __shfl_up(hi, 1, 8);
        const uint32_t lo = (k == 0) ? pr.first : prev;
        write_node(nodes, nkey, cb + k, pkey + (uint64_t)k * step, cl, lo, hi, g);
        flags[cb + k] = (hi - lo > max_leaf && cl < kMaxDepth) ? 1u : 0u;
    }
}

// k_recentre / k_node_ext over a node count read from the device (the level loop's total), launched
// with grids for the node capacity (surplus workgroups return at once: cheaper than grid-stride loops
// here, 0.55 + 0.15 against 0.59 + 0.17 ms at 2.06 M nodes of 3.1 M); k_recentre_dev also writes the
// offsets sentinel child[total] = 0.
__global__ __launch_bounds__(256) void k_recentre_dev(const double4 *__restrict__ p64, const NodeRec *__restrict__ nodes,
                                                     uint32_t *__restrict__ child, const uint32_t *__restrict__ total,
                                                     float4 *__restrict__ p32, float *__restrict__ amax) {
    const uint32_t n_total = *total;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) child[n_total] = 0u;
    const size_t u = t / kRecentreLanes;
    const uint32_t g = (uint32_t)(t % kRecentreLanes);
    if (u >= n_total) return;
    if (child[u] != 0) return;
    const float m = recentre_leaf<kRecentreLanes>(p64, nodes[u], g, p32);
    if (g == 0) amax[u] = m;
}

__global__ __launch_bounds__(256) void k_node_ext_dev(const NodeRec *__restrict__ nodes, const uint64_t *__restrict__ nkey,
                                                     const uint32_t *__restrict__ child, const float *__restrict__ amax,
                                                     const uint32_t *__restrict__ total, const LevelCtl *__restrict__ ctl,
                                                     int encoder, NodeExt *__restrict__ ext) {
    __shared__ LevelCtl s_ctl[kLvlCount];
    if (threadIdx.x < (uint32_t)kLvlCount) s_ctl[threadIdx.x] = ctl[threadIdx.x];
    __syncthreads();
    const uint32_t n_total = *total;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t u = t >> 3;
    if (u >= n_total) return;
    const uint32_t fc = child[u];
    if (fc == 0) return;
    uint32_t level = 0;
    for (uint32_t l = 0; l < (uint32_t)kLvlCount; l++)
        if (s_ctl[l].ls <= u && u < s_ctl[l].le) level = l;
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
    // node capacity: 8 n / M + 64 holds every tree seen (50 M tile, maxLeaf 128: 2.06 M nodes of 3.1 M);
    // the device level loop never grows it, the host loop below does when it must
    const size_t est = 64 + 8 * (size_t)n / (M ? M : 1);
    if ((e = grow_preserve(ctx, ctx->nodes, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->nkey, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->child, est + 1, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->flags, est, 0)) != hipSuccess) return e;
    if ((e = grow_preserve(ctx, ctx->nrank, est, 0)) != hipSuccess) return e;
    if ((e = ctx->misc.reserve(12 * 1024 + 64)) != hipSuccess) return e;
    uint32_t *total_dev = ctx->misc.p + 8;

    // leaf-recentred FP32 cloud and the children tables, sized for the capacity
    if ((e = ctx->p32.reserve(ctx->n)) != hipSuccess) return e;
    stage_begin(ctx, ST_BUILD);
    const size_t cap = std::min<size_t>(ctx->nodes.cap, ctx->child.cap - 1);
    if (cap < (1ull << 30)) {
        // device level loop, then the recentring and the children tables over the node count the loop
        // leaves on the device: one read-back at the end
        if ((e = ctx->amax.reserve(cap)) != hipSuccess) return e;
        if ((e = ctx->next.reserve((cap + 6) / 8 + 1)) != hipSuccess) return e;
        LevelCtl *ctl = reinterpret_cast<LevelCtl *>(ctx->misc.p + 4096);
        uint32_t *counters = ctx->misc.p + 4096 + 2 * (kLvlCount + 1), *errw = counters + kLvlCount, *totw = errw + 1;
        const size_t maxtiles = (cap + kLvlTile - 1) / kLvlTile, nlook = (size_t)kLvlCount * maxtiles;
        if ((e = ctx->lookback.reserve(nlook + 1)) != hipSuccess) return e;
        const unsigned ig = (unsigned)std::max<size_t>(1, std::min<size_t>((nlook + 255) / 256, 1024));
        k_level_init<<<ig, 256, 0, ctx->stream>>>(ctx->nodes.p, ctx->nkey.p, ctx->flags.p, n, M, g, ctl, ctx->lookback.p, nlook,
                                                  counters, errw, totw);
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
        const size_t sgrid = (size_t)std::max(1, cus) * 4, cgrid = (size_t)std::max(1, cus) * 8;
        for (int l = 0; l < kLvlFirstBatch; l++) {
            // level l holds at most min(8^l, cap) nodes: small grids for the top levels
            const size_t nmax = l < 10 ? std::min<size_t>(cap, 1ull << (3 * l)) : cap;
            const unsigned gs = (unsigned)std::min(sgrid, (nmax + kLvlTile - 1) / kLvlTile);
            const unsigned gc = (unsigned)std::min(cgrid, (nmax * 8 + 255) / 256);
            k_level_scan<<<gs, 256, 0, ctx->stream>>>(ctx->flags.p, ctx->child.p, ctl, (uint32_t)l,
                                                      ctx->lookback.p + (size_t)l * maxtiles, counters + l, (uint32_t)cap, errw, totw);
            k_level_children<<<gc, 256, 0, ctx->stream>>>(ctx->sorted_keys, ctx->nodes.p, ctx->nkey.p, ctx->child.p, ctx->flags.p,
                                                          ctl, (uint32_t)l, (uint32_t)cap, M, g, errw);
        }
        k_recentre_dev<<<(unsigned)((cap * kRecentreLanes + 255) / 256), 256, 0, ctx->stream>>>(
            ctx->p64.p, ctx->nodes.p, ctx->child.p, totw, ctx->p32.p, ctx->amax.p);
        k_node_ext_dev<<<(unsigned)((cap * 8 + 255) / 256), 256, 0, ctx->stream>>>(ctx->nodes.p, ctx->nkey.p, ctx->child.p,
                                                                                   ctx->amax.p, totw, ctl, ctx->encoder, ctx->next.p);
        stage_end(ctx, ST_BUILD);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        LevelCtl hc[kLvlCount + 1];
        uint32_t hw[2] = {0, 0};  // error word, node count
        if ((e = hipMemcpyAsync(hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(hw, errw, sizeof(hw), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
        if (hw[0] & kLvlErrSpin) {
            err = "build: look-back wait exceeded its bound";
            return hipErrorUnknown;
        }
        // complete within the first levels and the capacity: done (else the host loop below redoes it)
        if (!(hw[0] & kLvlErrCap) && hc[kLvlFirstBatch].le <= hc[kLvlFirstBatch].ls && hw[1] > 0) {
            uint32_t D = 0;  // the deepest non-empty level
            for (int l = 0; l < kLvlFirstBatch; l++)
                if (hc[l].le > hc[l].ls) D = (uint32_t)l;
            const uint32_t n_total = hc[D].le;
            if (n_total != hw[1]) {
                err = "build: inconsistent level table";
                return hipErrorUnknown;
            }
            for (uint32_t l = 0; l <= D; l++) ctx->level_start[l] = hc[l].ls;
            for (uint32_t l = D + 1; l < kMaxDepth + 3; l++) ctx->level_start[l] = n_total;
            ctx->n_total = n_total;
            ctx->n_internal = (n_total - 1) / 8;
            ctx->n_leaf = n_total - ctx->n_internal;
            ctx->n_levels = D + 1;
            ctx->max_depth_seen = D;
            ctx->max_points_leaf = M;
            return hipSuccess;
        }
    }

    // host level loop (arrays grown as needed): scan of the split flags, one round trip per level
    uint32_t ls = 0, le = 1, level = 0, n_internal = 0;
    k_root<<<1, 64, 0, ctx->stream>>>(ctx->nodes.p, ctx->nkey.p, ctx->flags.p, n, M, g);
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
