// Device code shared by the single-cloud build (build.hip) and the forest build of batches
// (pipeline.hip): child ranges by binary search of the sorted keys, node records, the leaf-recentred
// FP32 copy of a leaf's points, and the per-internal-node children table (NodeExt).
#pragma once

#include "device.hpp"

namespace oct {

// first index in [lo, hi) whose key is >= v (the reference's lower_bound counting,
// linear_octree.hpp:317-321)
__device__ inline uint32_t lower_bound_keys(const uint64_t *__restrict__ keys, uint32_t lo, uint32_t hi, uint64_t v) {
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// node u: start key, FP64 centre (getCenter, point_encoder.hpp:517-539) and point range
__device__ inline void put_node(NodeRec *__restrict__ nodes, uint64_t *__restrict__ nkey, uint32_t u, uint64_t key,
                                uint32_t level, uint32_t first, uint32_t last, int encoder, const double *boxmin,
                                const double *hl, const double *rad_level) {
    double c[3];
    node_center(encoder, key, level, boxmin, hl, rad_level, c);
    NodeRec r;
    r.cx = c[0], r.cy = c[1], r.cz = c[2];
    r.first = first, r.last = last;
    nodes[u] = r;
    nkey[u] = key;
}

// Leaf-recentred FP32 coordinates of one leaf, by a group of L lanes (L a power of two <= 64, the
// groups aligned in the wave; lane g of the group takes points first + g, first + g + L, ...):
// a = fl32(p - c_leaf) per axis (FP64 subtraction, one rounding to FP32), w = fl32(|a|^2) computed in
// FP64 from the FP32 values; returns (to every lane of the group) amax = max |a|^2 over the leaf
// rounded up, which bounds the FP32 rounding band, or +inf when a coordinate is not finite in FP32
// (the leaf's FP32 tests are then never trusted). The results do not depend on L.
template <int L>
__device__ inline float recentre_leaf(const double4 *__restrict__ p64, const NodeRec &r, uint32_t g,
                                      float4 *__restrict__ p32) {
    double m = 0.0;
    uint32_t bad = 0;
    for (uint32_t i = r.first + g; i < r.last; i += L) {
        double4 p = p64[i];
        float ax = (float)(p.x - r.cx), ay = (float)(p.y - r.cy), az = (float)(p.z - r.cz);
        double w = (double)ax * (double)ax + (double)ay * (double)ay + (double)az * (double)az;
        const float fw = (float)w;
        p32[i] = make_float4(ax, ay, az, fw);
        bad |= !(fw < __builtin_huge_valf()) ? 1u : 0u;
        m = w > m ? w : m;
    }
    for (int off = L / 2; off > 0; off >>= 1) {  // within the group: xor of a lane index below L
        const double o = __shfl_xor(m, off);
        m = o > m ? o : m;
        bad |= __shfl_xor(bad, off);
    }
    if (bad) m = __builtin_huge_val();
    return __double2float_ru(m);
}
constexpr int kRecentreLanes = 16;  // lanes per leaf (leaves hold at most maxLeaf points, ~30 on average)

// Children of internal node (first child fc, level `level`) regrouped by corner bits, one child per
// lane of a group of 8 (lane k takes child fc + k and writes its entries of the node's record *e).
// A child's bits come from its decoded key at level + 1, so they are exact also on degenerate axes
// where two children share a centre value; the 4 children on one side of an axis share that centre
// (the same corner coordinate), so their writes of it to e->ax are the same value.
__device__ inline void node_ext_child(const NodeRec *__restrict__ nodes, const uint64_t *__restrict__ nkey,
                                      const uint32_t *__restrict__ child, const float *__restrict__ amax, uint32_t fc,
                                      uint32_t k, uint32_t level, int encoder, NodeExt *__restrict__ e) {
    const uint32_t sh = kMaxDepth - (level + 1);  // corner bit of the children's level
    const uint32_t v = fc + k;
    uint32_t x, y, z;
    if (encoder == OCT_ENC_HILBERT3D)
        hilbert_decode(nkey[v], x, y, z);
    else
        morton_decode(nkey[v], x, y, z);
    const uint32_t bx = (x >> sh) & 1u, by = (y >> sh) & 1u, bz = (z >> sh) & 1u;
    const uint32_t c = bx | (by << 1) | (bz << 2);
    const NodeRec r = nodes[v];
    const uint32_t gc = child[v];
    e->ax[bx] = r.cx, e->ax[2 + by] = r.cy, e->ax[4 + bz] = r.cz;
    e->first[c] = r.first, e->last[c] = r.last;
    e->gc[c] = gc;
    e->amax[c] = gc == 0 ? amax[v] : 0.0f;
    if (k < 4) e->pad[k] = 0;
}

}  // namespace oct
