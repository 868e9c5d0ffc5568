// Device-driven build of a batch of clouds: bounding boxes, keys, a stable sort segmented by
// cloud, the sorted-cloud gather and a level-synchronous build of the whole forest, enqueued on one
// stream with no host synchronisation and no host-side sizes between the stages, so the sequence
// can be captured in a HIP graph and replayed.
//
// What each cloud gets is what the single-cloud path computes (and the reference):
//  - bbox: FP64 min / max per axis (box.hpp:94-137), centre = min + (max - min) / 2 and
//    radii = (max - min) / 2 (util.hpp:139-143, box.hpp:19);
//  - geometry: Box(centre, radii) corners, halfLengths = 0.5f * eps * (max' - min') and
//    radii[L] = 2^(21 - L) * halfLengths (linear_octree.hpp:439-455);
//  - keys: getAnchorCoords + Morton / Hilbert (point_encoder.hpp:37-49);
//  - order: stable sort of the cloud by key (point_encoder.hpp:285-383), here per cloud inside one
//    segmented LSD radix sort;
//  - tree: a node splits iff it holds more than maxPointsLeaf points and its level is below 21
//    (linear_octree.hpp:189), children ranges by lower_bound of the sorted keys (:317-321), centres
//    by getCenter (point_encoder.hpp:517-539).
// Forest layout: level 0 holds the C roots (root of cloud c at index c) padded with empty leaves to
// R = 1 (mod 8) slots; every later level holds all clouds' nodes of that level, cloud-major and
// key-minor, children of a node at level_start[l + 1] + 8 * (its rank among the split nodes of its
// level). Child groups of 8 therefore start at 1 (mod 8) and (first child - 1) / 8 indexes the
// per-internal-node table as in a single tree. Each cloud's nodes, taken in forest order, are its
// own tree in the reference's numbering (level-major, key-minor).
//
// Level sizes live in device memory (ctrl[kLv + l] = start of level l); each level step is three
// fixed-grid launches (block sums of the split flags, one-block scan of the sums, children), so a
// level that turns out empty costs three short launches and no host round trip. Node storage has a
// fixed capacity; a level that would overflow it raises an error bit and adds no children, and the
// host grows the capacity and runs the build again.
#include <cfloat>

#include "tree_common.hpp"
#include "pipeline.hpp"

namespace oct {

constexpr int kLevelBlocks = 512;  // fixed grid of the level steps
constexpr int kLevelThreads = 256;

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ bbox per chunk, geometry per cloud
__device__ inline void minmax_combine(double *mn, double *mx, const double *omn, const double *omx) {
    for (int a = 0; a < 3; a++) {
        mn[a] = (omn[a] < mn[a]) ? omn[a] : mn[a];  // std::min semantics (NaN never replaces the running value)
        mx[a] = (mx[a] < omx[a]) ? omx[a] : mx[a];
    }
}

__global__ __launch_bounds__(256) void k_bp_bbox(const double *__restrict__ xyz, const Chunk *__restrict__ chunks,
                                                 double *__restrict__ partial) {
    const Chunk ch = chunks[blockIdx.x];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    const size_t end = (size_t)ch.start + ch.len;
    for (size_t i = (size_t)ch.start + threadIdx.x; i < end; i += blockDim.x) {
        double v[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        minmax_combine(mn, mx, v, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        double omn[3], omx[3];
        for (int a = 0; a < 3; a++) omn[a] = __shfl_xor(mn[a], off), omx[a] = __shfl_xor(mx[a], off);
        minmax_combine(mn, mx, omn, omx);
    }
    __shared__ double s[4][6];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int a = 0; a < 3; a++) s[w][a] = mn[a], s[w][3 + a] = mx[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < 4; j++) minmax_combine(mn, mx, s[j], s[j] + 3);
        for (int a = 0; a < 3; a++) partial[(size_t)blockIdx.x * 6 + a] = mn[a], partial[(size_t)blockIdx.x * 6 + 3 + a] = mx[a];
    }
}

// One block per cloud: its chunks' partial boxes, then the cloud's Box and octree geometry. An empty
// cloud keeps min = DBL_MAX, max = -DBL_MAX, as the single-cloud path.
__global__ __launch_bounds__(256) void k_bp_geometry(const double *__restrict__ partial, const uint32_t *__restrict__ cloud_chunk0,
                                                     const uint32_t *__restrict__ cloud_enc,
                                                     const uint32_t *__restrict__ cloud_leaf, CloudGeom *__restrict__ geom,
                                                     double *__restrict__ radtab, double *__restrict__ magtab) {
    const uint32_t c = blockIdx.x;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t k = cloud_chunk0[c] + threadIdx.x; k < cloud_chunk0[c + 1]; k += blockDim.x)
        minmax_combine(mn, mx, partial + (size_t)k * 6, partial + (size_t)k * 6 + 3);
    for (int off = 32; off > 0; off >>= 1) {
        double omn[3], omx[3];
        for (int a = 0; a < 3; a++) omn[a] = __shfl_xor(mn[a], off), omx[a] = __shfl_xor(mx[a], off);
        minmax_combine(mn, mx, omn, omx);
    }
    __shared__ double s[4][6];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int a = 0; a < 3; a++) s[w][a] = mn[a], s[w][3 + a] = mx[a];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int j = 1; j < 4; j++) minmax_combine(mn, mx, s[j], s[j] + 3);
    CloudGeom g;
    const double eps = (double)(1.0f / (float)(1 << kMaxDepth));
    for (int a = 0; a < 3; a++) {
        g.center[a] = mn[a] + (mx[a] - mn[a]) / 2;
        g.radii[a] = (mx[a] - mn[a]) / 2;
        const double lo = g.center[a] - g.radii[a], hi = g.center[a] + g.radii[a];
        g.boxmin[a] = lo;
        g.hl[a] = (double)0.5f * eps * (hi - lo);
    }
    for (uint32_t l = 0; l <= kMaxDepth; l++) {
        const double side = (double)(1u << (kMaxDepth - l));
        for (int a = 0; a < 3; a++) {
            g.rad[l][a] = side * g.hl[a];
            radtab[(size_t)c * 66 + 3 * l + a] = g.rad[l][a];
        }
    }
    for (int a = 0; a < 3; a++) magtab[(size_t)c * 3 + a] = fabs(g.center[a]) + fabs(g.radii[a]);
    g.encoder = cloud_enc[c];
    g.max_leaf = cloud_leaf[c];
    g.pad[0] = g.pad[1] = 0;
    geom[c] = g;
}

// ------------------------------------------------------------------ keys + per-cloud digit histograms
__global__ __launch_bounds__(256) void k_bp_encode(const double *__restrict__ xyz, const Chunk *__restrict__ chunks,
                                                   const CloudGeom *__restrict__ geom, uint64_t *__restrict__ keys,
                                                   uint32_t *__restrict__ hist) {
    __shared__ uint32_t sh[kRadixPassCount * 256];
    __shared__ uint16_t stab[48 * 8];
    __shared__ uint16_t stab2[kHilbertTab2];
    const Chunk ch = chunks[blockIdx.x];
    const int enc = (int)geom[ch.cloud].encoder;
    double c[3], r[3];
    for (int a = 0; a < 3; a++) c[a] = geom[ch.cloud].center[a], r[a] = geom[ch.cloud].radii[a];
    for (int i = threadIdx.x; i < kRadixPassCount * 256; i += blockDim.x) sh[i] = 0;
    for (int i = threadIdx.x; i < 48 * 8; i += blockDim.x) stab[i] = kHilbert.t[i];
    __syncthreads();
    if (enc == OCT_ENC_HILBERT3D) {  // the same for the whole block
        hilbert_tab2_fill(stab2, stab, threadIdx.x, blockDim.x);
        __syncthreads();
    }
    const size_t end = (size_t)ch.start + ch.len;
    for (size_t i = (size_t)ch.start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t x = quantize_axis(xyz[3 * i + 0], c[0], r[0]);
        const uint32_t y = quantize_axis(xyz[3 * i + 1], c[1], r[1]);
        const uint32_t z = quantize_axis(xyz[3 * i + 2], c[2], r[2]);
        const uint64_t k = (enc == OCT_ENC_HILBERT3D) ? hilbert_encode_tab2(stab, stab2, x, y, z) : morton_encode(x, y, z);
        keys[i] = k;
#pragma unroll
        for (int p = 0; p < kRadixPassCount; p++) atomicAdd(&sh[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    uint32_t *h = hist + (size_t)ch.cloud * kSegBaseStride;
    for (int i = threadIdx.x; i < kRadixPassCount * 256; i += blockDim.x)
        if (sh[i]) atomicAdd(&h[i], sh[i]);
}

// Per (cloud, pass): bases[cloud][pass][d] = cloud offset + number of the cloud's keys with a smaller
// digit. One block per cloud, one wave per pass, 4 digits per lane.
__global__ __launch_bounds__(512) void k_bp_bases(const uint32_t *__restrict__ hist, const uint32_t *__restrict__ cloud_off,
                                                  uint32_t *__restrict__ bases) {
    const uint32_t c = blockIdx.x, p = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t *h = hist + (size_t)c * kSegBaseStride + p * 256 + lane * 4;
    const uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3];
    const uint32_t sum = v0 + v1 + v2 + v3;
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += o;
    }
    uint32_t run = cloud_off[c] + incl - sum;
    uint32_t *o = bases + (size_t)c * kSegBaseStride + p * 256 + lane * 4;
    o[0] = run, run += v0;
    o[1] = run, run += v1;
    o[2] = run, run += v2;
    o[3] = run;
}

// ------------------------------------------------------------------ forest build
// Level 0: root of cloud r at index r (key 0, the cloud's whole range), empty padding leaves up to R.
__global__ void k_bp_roots(const CloudGeom *__restrict__ geom, const uint32_t *__restrict__ cloud_off, uint32_t C, uint32_t R,
                           NodeRec *__restrict__ nodes, uint64_t *__restrict__ nkey, uint32_t *__restrict__ flags,
                           uint32_t *__restrict__ ncloud, uint32_t *__restrict__ roots, uint32_t *__restrict__ lv) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) {
        lv[0] = 0;
        for (uint32_t l = 1; l < kLvCount; l++) lv[l] = R;
    }
    if (r >= R) return;
    if (r < C) {
        const CloudGeom &g = geom[r];
        const uint32_t first = cloud_off[r], last = cloud_off[r + 1];
        put_node(nodes, nkey, r, 0ull, 0, first, last, (int)g.encoder, g.boxmin, g.hl, g.rad[0]);
        flags[r] = (last - first > g.max_leaf) ? 1u : 0u;
        ncloud[r] = r;
        roots[r] = r;
    } else {
        NodeRec z;
        z.cx = z.cy = z.cz = 0.0;
        z.first = z.last = 0;
        nodes[r] = z;
        nkey[r] = 0ull;
        flags[r] = 0u;
        ncloud[r] = kNoCloud;
    }
}

// the slice [b0, b1) of level l's nodes that block `blk` of the fixed level grid owns
__device__ inline void level_slice(const uint32_t *lv, uint32_t l, uint32_t blk, uint32_t &b0, uint32_t &b1) {
    const uint32_t ls = lv[l], le = lv[l + 1];
    const uint32_t cnt = le - ls;
    const uint32_t chunk = (cnt + kLevelBlocks - 1) / kLevelBlocks;
    const uint64_t s = (uint64_t)ls + (uint64_t)blk * chunk;
    b0 = (uint32_t)(s < le ? s : le);
    b1 = (uint32_t)((uint64_t)b0 + chunk < le ? b0 + chunk : le);
}

__device__ inline uint32_t block_sum_u32(uint32_t v, uint32_t *s_w) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) s_w[w] = v;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t k = 0; k < blockDim.x / 64; k++) t += s_w[k];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kLevelThreads) void k_bp_level_sum(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ lv,
                                                                uint32_t l, uint32_t *__restrict__ part) {
    __shared__ uint32_t s_w[kLevelThreads / 64];
    uint32_t b0, b1;
    level_slice(lv, l, blockIdx.x, b0, b1);
    uint32_t v = 0;
    for (uint32_t u = b0 + threadIdx.x; u < b1; u += blockDim.x) v += flags[u];
    const uint32_t t = block_sum_u32(v, s_w);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// One block: exclusive scan of the block sums in place; start of level l + 2 = end of level l + 1.
__global__ __launch_bounds__(kLevelBlocks) void k_bp_level_scan(uint32_t *__restrict__ part, uint32_t *__restrict__ lv, uint32_t l,
                                                                uint32_t cap, uint32_t *__restrict__ build_err) {
    __shared__ uint32_t s_w[kLevelBlocks / 64];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    const uint32_t v = part[t];
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += o;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t pre = 0, total = 0;
    for (uint32_t k = 0; k < kLevelBlocks / 64; k++) {
        if (k < w) pre += s_w[k];
        total += s_w[k];
    }
    part[t] = pre + incl - v;
    if (t == 0) {
        const uint32_t le = lv[l + 1];
        uint64_t next = (uint64_t)le + 8ull * total;
        if (total && l >= kMaxDepth) {
            atomicOr(build_err, kBuildErrDepth);
            next = le;
        }
        if (next > cap) {
            atomicOr(build_err, kBuildErrCapacity);
            next = le;
        }
        for (uint32_t k = l + 2; k < kLvCount; k++) lv[k] = (uint32_t)next;
    }
}

// Children of the split nodes of level l: block `blk` walks its slice in rounds of 256 nodes, ranks
// the split flags (wave ballots + a block scan) on top of its exclusive block offset, and writes each
// split node's 8 children (range, key, centre, flag, cloud) at level_start[l + 1] + 8 * rank, one
// lane per child.
__global__ __launch_bounds__(kLevelThreads) void k_bp_level_children(
    const uint64_t *__restrict__ keys, NodeRec *__restrict__ nodes, uint64_t *__restrict__ nkey, uint32_t *__restrict__ child,
    uint32_t *__restrict__ flags, uint32_t *__restrict__ ncloud, const CloudGeom *__restrict__ geom,
    const uint32_t *__restrict__ lv, uint32_t l, const uint32_t *__restrict__ part, uint32_t *__restrict__ cl_internal,
    uint32_t *__restrict__ cl_depth) {
    __shared__ uint32_t s_w[kLevelThreads / 64];
    __shared__ uint32_t s_u[kLevelThreads], s_cb[kLevelThreads], s_cid[kLevelThreads];  // a round's split nodes
    uint32_t b0, b1;
    level_slice(lv, l, blockIdx.x, b0, b1);
    const uint32_t le = lv[l + 1];
    const bool allow = lv[l + 2] != le;  // false: no children at this level (none split, or an error)
    uint32_t running = part[blockIdx.x];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t cl = l + 1;
    for (uint32_t base = b0; base < b1; base += blockDim.x) {
        const uint32_t u = base + threadIdx.x;
        const bool in = u < b1;
        const uint32_t f = (in && allow) ? flags[u] : 0u;
        const uint64_t m = ballot64(f != 0);
        const uint32_t below = (uint32_t)__popcll(m & lanemask_lt());
        if (lane == 0) s_w[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (uint32_t k = 0; k < kLevelThreads / 64; k++) {
            if (k < w) wpre += s_w[k];
            tot += s_w[k];
        }
        __syncthreads();
        uint32_t cid = 0;
        if (in) {
            if (!f) {
                child[u] = 0;
            } else {
                cid = ncloud[u];
                const uint32_t j = wpre + below;  // the split node's slot in this round
                const uint32_t cb = le + 8u * (running + j);
                child[u] = cb;
                s_u[j] = u, s_cb[j] = cb, s_cid[j] = cid;
            }
        }
        __syncthreads();
        // the round's split nodes, 8 lanes each: lane ck finds child ck's end and takes its start from
        // lane ck - 1 (a search over the whole parent range finds the same index as one from the start)
        for (uint32_t q = threadIdx.x >> 3; q < tot; q += blockDim.x >> 3) {
            const uint32_t ck = threadIdx.x & 7u;
            const uint32_t pu = s_u[q], cb = s_cb[q], cc = s_cid[q];
            const CloudGeom &g = geom[cc];
            const NodeRec pr = nodes[pu];
            const uint64_t pkey = nkey[pu];
            const uint64_t step = node_range(cl);
            const uint32_t hi = (ck == 7) ? pr.last : lower_bound_keys(keys, pr.first, pr.last, pkey + (uint64_t)(ck + 1) * step);
            const uint32_t prev = __shfl_up(hi, 1, 8);
            const uint32_t lo = (ck == 0) ? pr.first : prev;
            put_node(nodes, nkey, cb + ck, pkey + (uint64_t)ck * step, cl, lo, hi, (int)g.encoder, g.boxmin, g.hl, g.rad[cl]);
            flags[cb + ck] = (hi - lo > g.max_leaf && cl < kMaxDepth) ? 1u : 0u;
            ncloud[cb + ck] = cc;
        }
        // per-cloud internal-node counts and depth: one atomic per wave when its split nodes share a cloud
        if (m) {
            const uint32_t lead = (uint32_t)__builtin_ctzll(m);
            const uint32_t c0 = readlane_u32(cid, (int)lead);
            const bool same = !f || cid == c0;
            if (ballot64(!same) == 0) {
                if (lane == lead) {
                    atomicAdd(&cl_internal[c0], (uint32_t)__popcll(m));
                    atomicMax(&cl_depth[c0], cl);
                }
            } else if (f) {
                atomicAdd(&cl_internal[cid], 1u);
                atomicMax(&cl_depth[cid], cl);
            }
        }
        running += tot;
        __syncthreads();  // the next round rewrites the slot lists
    }
}

// sentinel child entry (offsets has nTotal + 1 entries in the reference, linear_octree.hpp:497)
__global__ void k_bp_sentinel(const uint32_t *__restrict__ lv, uint32_t *__restrict__ child) {
    if (blockIdx.x == 0 && threadIdx.x == 0) child[lv[kLvCount - 1]] = 0;
}

// kRecentreLanes lanes per leaf, grid-stride over the forest's nodes
__global__ __launch_bounds__(256) void k_bp_recentre(const double4 *__restrict__ p64, const NodeRec *__restrict__ nodes,
                                                     const uint32_t *__restrict__ child, const uint32_t *__restrict__ lv,
                                                     float4 *__restrict__ p32, float *__restrict__ amax) {
    const uint32_t n_total = lv[kLvCount - 1];
    const uint32_t g = threadIdx.x % kRecentreLanes;
    const uint32_t groups = (gridDim.x * blockDim.x) / kRecentreLanes;
    for (uint32_t u = (blockIdx.x * blockDim.x + threadIdx.x) / kRecentreLanes; u < n_total; u += groups) {
        if (child[u] != 0) continue;  // the same for the whole group
        const float m = recentre_leaf<kRecentreLanes>(p64, nodes[u], g, p32);
        if (g == 0) amax[u] = m;
    }
}

__global__ __launch_bounds__(256) void k_bp_node_ext(const NodeRec *__restrict__ nodes, const uint64_t *__restrict__ nkey,
                                                     const uint32_t *__restrict__ child, const float *__restrict__ amax,
                                                     const uint32_t *__restrict__ lv, const uint32_t *__restrict__ ncloud,
                                                     const CloudGeom *__restrict__ geom, NodeExt *__restrict__ ext) {
    const uint32_t n_total = lv[kLvCount - 1];
    const uint32_t k = threadIdx.x & 7u;  // child of the node: 8 lanes per node
    const uint32_t groups = (gridDim.x * blockDim.x) >> 3;
    for (uint32_t u = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; u < n_total; u += groups) {
        const uint32_t fc = child[u];
        if (fc == 0) continue;
        uint32_t level = 0;
        while (level + 1 < kLvCount && lv[level + 1] <= u) level++;
        node_ext_child(nodes, nkey, child, amax, fc, k, level, (int)geom[ncloud[u]].encoder, ext + ((fc - 1) >> 3));
    }
}

// ------------------------------------------------------------------ query grouping
// sort keys (class << 32 | global centre index), global centre indices, digit histograms
__global__ __launch_bounds__(256) void k_q_keys(const uint32_t *__restrict__ cloud, const uint32_t *__restrict__ idx,
                                                const uint32_t *__restrict__ kq, const uint32_t *__restrict__ cloud_off,
                                                uint32_t n, uint64_t *__restrict__ keys, uint32_t *__restrict__ gidx,
                                                uint32_t *__restrict__ hist) {
    __shared__ uint32_t sh[5 * 256];
    for (int i = threadIdx.x; i < 5 * 256; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const int np = kq ? 5 : 4;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = cloud_off[cloud[i]] + idx[i];
        const uint64_t key = (uint64_t)g | (kq ? (uint64_t)knn_class_of(kq[i]) << 32 : 0ull);
        keys[i] = key;
        gidx[i] = g;
        for (int p = 0; p < np; p++) atomicAdd(&sh[p * 256 + ((key >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np * 256; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// first sorted position of every (class, cloud) group; gstart[nclass * C] = n
__global__ void k_q_group_starts(const uint64_t *__restrict__ skeys, uint32_t n, const uint32_t *__restrict__ cloud_off,
                                 uint32_t C, uint32_t nclass, uint32_t *__restrict__ gstart) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > nclass * C) return;
    if (g == nclass * C) {
        gstart[g] = n;
        return;
    }
    const uint64_t v = ((uint64_t)(g / C) << 32) | cloud_off[g % C];
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (skeys[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    gstart[g] = lo;
}

struct ClassShape {
    uint32_t G[kKnnClasses];
    uint32_t slots[kKnnClasses];
    uint32_t *order[kKnnClasses];
    uint32_t *tile_info[kKnnClasses];
};

// one block per class: wave tiles per group, exclusive offsets over the class's groups, tile count
__global__ __launch_bounds__(1024) void k_q_tile_offsets(const uint32_t *__restrict__ gstart, uint32_t C, ClassShape sh,
                                                         uint32_t *__restrict__ toff, uint32_t *__restrict__ ntiles) {
    __shared__ uint32_t s_w[16];
    const uint32_t k = blockIdx.x, t = threadIdx.x, w = t >> 6, lane = t & 63;
    const uint32_t G = sh.G[k];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < C; base += 1024) {
        const uint32_t c = base + t;
        uint32_t nt = 0;
        if (c < C) {
            const uint32_t g = k * C + c;
            nt = (gstart[g + 1] - gstart[g] + G - 1) / G;
        }
        uint32_t incl = nt;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off);
            if (lane >= (uint32_t)off) incl += o;
        }
        if (lane == 63) s_w[w] = incl;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (uint32_t j = 0; j < 16; j++) {
            if (j < w) pre += s_w[j];
            tot += s_w[j];
        }
        if (c < C) toff[k * C + c] = carry + pre + incl - nt;
        carry += tot;
        __syncthreads();
    }
    if (t == 0) ntiles[k] = carry;
}

// slot s of class blockIdx.y: its tile t = s / G, the group holding t (last group whose offset is
// <= t), the query at that group's position; tile_info written by the tile's first slot
__global__ __launch_bounds__(256) void k_q_tiles(const uint32_t *__restrict__ svals, const uint32_t *__restrict__ gstart,
                                                 const uint32_t *__restrict__ toff, const uint32_t *__restrict__ ntiles,
                                                 uint32_t C, ClassShape sh) {
    const uint32_t k = blockIdx.y;
    const uint32_t G = sh.G[k];
    const uint32_t nt = ntiles[k];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < sh.slots[k]; s += gridDim.x * blockDim.x) {
        const uint32_t t = s / G, j = s % G;
        if (t >= nt) break;
        const uint32_t *to = toff + (size_t)k * C;
        uint32_t lo = 0, hi = C;  // first group offset > t
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (to[mid] <= t)
                lo = mid + 1;
            else
                hi = mid;
        }
        const uint32_t c = lo - 1, g = k * C + c;
        const uint32_t start = gstart[g] + (t - to[c]) * G;
        const uint32_t left = gstart[g + 1] - start;
        const uint32_t nvalid = left < G ? left : G;
        sh.order[k][s] = j < nvalid ? svals[start + j] : 0u;
        if (j == 0) sh.tile_info[k][t] = c | (nvalid << 20);
    }
}

static inline unsigned grid_for(size_t n, unsigned per, unsigned cap) {
    size_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return (unsigned)(b > cap ? cap : b);
}

hipError_t enqueue_forest_build(hipStream_t s, const ForestBuild &F) {
    // clear what accumulates: sort tile counters and flags, level table, error bits, histograms,
    // look-back words, per-cloud counters (fill kernels, which graph replays run like any launch)
    k_fill_u32<<<1, 256, 0, s>>>(F.ctrl, kCtrlWords, 0u);
    k_fill_u32<<<grid_for((size_t)F.C * kSegBaseStride, 256, 4096), 256, 0, s>>>(F.hist, (size_t)F.C * kSegBaseStride, 0u);
    const size_t lbw = (size_t)kRadixPassCount * F.ntiles * 256;
    k_fill_u32<<<grid_for(lbw, 256, 8192), 256, 0, s>>>(F.lookback, lbw, 0u);
    k_fill_u32<<<grid_for(F.C, 256, 64), 256, 0, s>>>(F.cl_internal, F.C, 0u);
    k_fill_u32<<<grid_for(F.C, 256, 64), 256, 0, s>>>(F.cl_depth, F.C, 0u);
    if (F.nchunks) k_bp_bbox<<<F.nchunks, 256, 0, s>>>(F.xyz, F.chunks, F.partial);
    k_bp_geometry<<<F.C, 256, 0, s>>>(F.partial, F.cloud_chunk0, F.cloud_enc, F.cloud_leaf, F.geom, F.radtab, F.magtab);
    if (F.nchunks) k_bp_encode<<<F.nchunks, 256, 0, s>>>(F.xyz, F.chunks, F.geom, F.keys[0], F.hist);
    k_bp_bases<<<F.C, 512, 0, s>>>(F.hist, F.cloud_off, F.bases);
    hipError_t e = launch_sort_segmented(s, F.keys[0], F.keys[1], F.vals[0], F.vals[1], F.n, F.tiles, F.ntiles, F.bases,
                                         F.lookback, F.ctrl + kCtrlSortCounters, F.ctrl + kCtrlSortErr);
    if (e != hipSuccess) return e;
    if ((e = launch_gather_raw(s, F.xyz, F.vals[0], F.p64, F.n)) != hipSuccess) return e;
    uint32_t *lv = F.ctrl + kCtrlLv;
    k_bp_roots<<<grid_for(F.R, 256, 1u << 16), 256, 0, s>>>(F.geom, F.cloud_off, F.C, F.R, F.nodes, F.nkey, F.flags, F.ncloud,
                                                            F.roots, lv);
    for (uint32_t l = 0; l <= kMaxDepth; l++) {
        k_bp_level_sum<<<kLevelBlocks, kLevelThreads, 0, s>>>(F.flags, lv, l, F.part);
        k_bp_level_scan<<<1, kLevelBlocks, 0, s>>>(F.part, lv, l, F.cap, F.ctrl + kCtrlBuildErr);
        k_bp_level_children<<<kLevelBlocks, kLevelThreads, 0, s>>>(F.keys[0], F.nodes, F.nkey, F.child, F.flags, F.ncloud, F.geom,
                                                                   lv, l, F.part, F.cl_internal, F.cl_depth);
    }
    k_bp_sentinel<<<1, 64, 0, s>>>(lv, F.child);
    k_bp_recentre<<<4096, 256, 0, s>>>(F.p64, F.nodes, F.child, lv, F.p32, F.amax);
    k_bp_node_ext<<<2048, 256, 0, s>>>(F.nodes, F.nkey, F.child, F.amax, lv, F.ncloud, F.geom, F.ext);
    return hipGetLastError();
}

hipError_t enqueue_query_groups(hipStream_t s, const QueryGroups &Q, uint32_t C, const uint32_t *cloud_off) {
    k_fill_u32<<<1, 256, 0, s>>>(Q.ctrl, kCtrlWords, 0u);
    k_fill_u32<<<1, 256, 0, s>>>(Q.hist, (size_t)kRadixPassCount * 256, 0u);
    if (Q.n == 0) return hipGetLastError();  // tile counts stay 0: the work queues end at once
    const int np = Q.kq ? 5 : 4;
    const size_t tiles = (Q.n + kSortTile - 1) / kSortTile;
    const size_t lbw = (size_t)np * tiles * 256;
    k_fill_u32<<<grid_for(lbw, 256, 8192), 256, 0, s>>>(Q.lookback, lbw, 0u);
    k_q_keys<<<grid_for(Q.n, 256, 2048), 256, 0, s>>>(Q.cloud, Q.idx, Q.kq, cloud_off, Q.n, Q.keys[0], Q.gidx, Q.hist);
    k_bp_bases<<<1, 512, 0, s>>>(Q.hist, Q.ctrl + kQCtrlZero, Q.bases);
    hipError_t e = launch_sort_passes(s, Q.keys[0], Q.keys[1], Q.vals[0], Q.vals[1], Q.n, np, Q.bases, Q.lookback,
                                      Q.ctrl + kCtrlSortCounters, Q.ctrl + kCtrlSortErr);
    if (e != hipSuccess) return e;
    const uint64_t *sk = Q.keys[np & 1];
    const uint32_t *sv = Q.vals[np & 1];
    const uint32_t ng = Q.nclass * C + 1;
    k_q_group_starts<<<grid_for(ng, 256, 1u << 20), 256, 0, s>>>(sk, Q.n, cloud_off, C, Q.nclass, Q.gstart);
    ClassShape sh;
    uint32_t maxslots = 0;
    for (int k = 0; k < kKnnClasses; k++) {
        sh.G[k] = Q.G[k], sh.slots[k] = (uint32_t)k < Q.nclass ? Q.slots[k] : 0u;
        sh.order[k] = Q.order[k], sh.tile_info[k] = Q.tile_info[k];
        maxslots = sh.slots[k] > maxslots ? sh.slots[k] : maxslots;
    }
    k_q_tile_offsets<<<Q.nclass, 1024, 0, s>>>(Q.gstart, C, sh, Q.toff, Q.ctrl + kQCtrlTiles);
    k_q_tiles<<<dim3(grid_for(maxslots, 256, 4096), Q.nclass), 256, 0, s>>>(sv, Q.gstart, Q.toff, Q.ctrl + kQCtrlTiles, C, sh);
    return hipGetLastError();
}

}  // namespace oct
