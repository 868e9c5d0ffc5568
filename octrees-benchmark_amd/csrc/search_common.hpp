// Declarations of the radius-search kernels (search.hip): launch
// arguments, per-query state and the reference's FP64 predicates.
#pragma once

#include "device.hpp"

namespace oct {

constexpr int kSearchBlock = 256;
constexpr int kWaves = kSearchBlock / 64;
constexpr int kStack = 7 * 21 + 13;  // depth first over <= 21 internal levels: <= 7 pending siblings per level + 8

enum { J_OUTSIDE = 0, J_OVERLAP = 1, J_INSIDE = 2 };

struct SearchArgs {
    const double4 *__restrict__ p64;
    const float4 *__restrict__ p32;
    const float *__restrict__ amax;
    const NodeRec *__restrict__ nodes;
    const NodeExt *__restrict__ next;
    const uint32_t *__restrict__ child;
    const uint32_t *__restrict__ order;
    const uint32_t *__restrict__ qidx;
    const double *__restrict__ qxyz;
    const double *__restrict__ qrad;
    double radius;
    const double *__restrict__ qrad3;
    double r3[3];
    int use_r3;
    int remap;  // count kernel: compact the queries of a leaf wanted by at most half the wave
    // neighborsApprox (algorithm OCT_ALGO_APPROX: Struct with a depth cutoff): nodes deeper than
    // approx_level are added whole (approx_upper) or dropped without a kernel test
    uint32_t approx_level;
    int approx_upper;
    // Count kernel's tree geometry and roots: level radii radtab[cloud * 66 + 3 * level + axis]; for a
    // forest of clouds (batches), tile_info[wave tile] = cloud | valid slots << 20 and roots[cloud]
    // (null: one cloud, root node 0, every slot below nq valid).
    const double *__restrict__ radtab;
    const uint32_t *__restrict__ tile_info;
    const uint32_t *__restrict__ roots;
    // work-queue form of the count kernel: number of wave tiles (device-resident) and the counter the
    // waves take tiles from (zero at launch); null queue_next: one wave tile per wave of the grid
    const uint32_t *__restrict__ queue_tiles;
    uint32_t *__restrict__ queue_next;
    uint32_t nq;
    uint32_t *__restrict__ counts;
    uint32_t *__restrict__ ecounts;
    uint64_t *__restrict__ ecounts64;  // list sizes pass: entries per query as u64 (scanned in place)
    const uint64_t *__restrict__ offsets;
    uint32_t *__restrict__ entries;
    unsigned long long *__restrict__ stats;
    uint32_t *__restrict__ err;  // device error word: kErrStack when a traversal stack was full
    double rad[22][3];
};

struct Query {
    double x, y, z, r, R;
    double lox, loy, loz, hix, hiy, hiz;  // kernel box: center -/+ radius (kernel_abstract.hpp:41-45)
};

// Read-only tree records through the constant address space: a wave-uniform index then compiles
// to scalar loads (the records land in SGPRs, not in 64 copies across the vector registers).
typedef const __attribute__((address_space(4))) NodeExt ConstNodeExt;
typedef const __attribute__((address_space(4))) double ConstDouble;
__device__ inline ConstDouble *const_doubles(const double *p) { return (ConstDouble *)(uintptr_t)p; }
__device__ inline ConstNodeExt &node_ext(const NodeExt *p, uint32_t i) {
    return ((ConstNodeExt *)(uintptr_t)p)[i];
}

__device__ inline double dmax0(double a) { return (a < 0.0) ? 0.0 : a; }  // std::max(a, 0.0)

// Cube / square kernels built from a Vector of half extents (KernelAbstract(center, radii) and
// makeBox, kernel_abstract.hpp:39-45): box = centre -/+ radii per axis.
__device__ inline void box_radii(const SearchArgs &a, bool valid, uint32_t qid, Query &q) {
    if (!a.use_r3) return;
    double rx = a.r3[0], ry = a.r3[1], rz = a.r3[2];
    if (a.qrad3 && valid) rx = a.qrad3[3 * (size_t)qid], ry = a.qrad3[3 * (size_t)qid + 1], rz = a.qrad3[3 * (size_t)qid + 2];
    q.lox = q.x - rx, q.loy = q.y - ry, q.loz = q.z - rz;
    q.hix = q.x + rx, q.hiy = q.y + ry, q.hiz = q.z + rz;
}

// boxIntersect(center, Vector radii) of the four kernels (kernel_3d.hpp:181-223, 277-302;
// kernel_2d.hpp:107-144, 195-218)
template <int KERNEL>
__device__ inline int classify64(const Query &q, double cx, double cy, double cz, double rx, double ry, double rz) {
    if (KERNEL == OCT_KERNEL_SPHERE) {
        double dx = fabs(q.x - cx), dy = fabs(q.y - cy), dz = fabs(q.z - cz);
        const double mdx = q.r + rx, mdy = q.r + ry, mdz = q.r + rz;
        if (dx > mdx || dy > mdy || dz > mdz) return J_OUTSIDE;
        double ex = dx + rx, ey = dy + ry, ez = dz + rz;
        if ((ex * ex + ey * ey + ez * ez) <= q.R) return J_INSIDE;
        int ni = (dx < rx) + (dy < ry) + (dz < rz);
        if (ni > 1) return J_OVERLAP;
        dx = dmax0(dx - rx), dy = dmax0(dy - ry), dz = dmax0(dz - rz);
        if ((dx * dx + dy * dy + dz * dz) <= q.R) return J_OVERLAP;
        return J_OUTSIDE;
    } else if (KERNEL == OCT_KERNEL_CIRCLE) {
        double dx = fabs(q.x - cx), dy = fabs(q.y - cy);
        const double mdx = q.r + rx, mdy = q.r + ry;
        if (dx > mdx || dy > mdy) return J_OUTSIDE;
        double ex = dx + rx, ey = dy + ry;
        if ((ex * ex + ey * ey) <= q.R) return J_INSIDE;
        int ni = (dx < rx) + (dy < ry);
        if (ni > 1) return J_OVERLAP;
        dx = dmax0(dx - rx), dy = dmax0(dy - ry);
        if ((dx * dx + dy * dy) <= q.R) return J_OVERLAP;
        return J_OUTSIDE;
    } else if (KERNEL == OCT_KERNEL_CUBE) {
        const double hx = cx + rx, lx = cx - rx, hy = cy + ry, ly = cy - ry, hz = cz + rz, lz = cz - rz;
        if (hx < q.lox || hy < q.loy || hz < q.loz || lx > q.hix || ly > q.hiy || lz > q.hiz) return J_OUTSIDE;
        if (hx <= q.hix && hy <= q.hiy && hz <= q.hiz && lx >= q.lox && ly >= q.loy && lz >= q.loz) return J_INSIDE;
        return J_OVERLAP;
    } else {
        const double hx = cx + rx, lx = cx - rx, hy = cy + ry, ly = cy - ry;
        if (hx < q.lox || hy < q.loy || lx > q.hix || ly > q.hiy) return J_OUTSIDE;
        if (hx <= q.hix && hy <= q.hiy && lx >= q.lox && ly >= q.loy) return J_INSIDE;
        return J_OVERLAP;
    }
}

// boxOverlap(center, Vector radii): Kernel3D (kernel_3d.hpp:100-115) / Kernel2D (kernel_2d.hpp:29-41)
template <int KERNEL>
__device__ inline bool overlap64(const Query &q, double cx, double cy, double cz, double rx, double ry, double rz) {
    constexpr bool is3d = (KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CUBE);
    if (cx + rx < q.lox || cy + ry < q.loy || (is3d && cz + rz < q.loz)) return false;
    if (cx - rx > q.hix || cy - ry > q.hiy || (is3d && cz - rz > q.hiz)) return false;
    return true;
}

// isInside (kernel_3d.hpp:133-138, 238-242; kernel_2d.hpp:59-63, 159-162; util.hpp:133-137)
template <int KERNEL>
__device__ inline bool inside64(const Query &q, double px, double py, double pz) {
    if (KERNEL == OCT_KERNEL_SPHERE) {
        double d1 = px - q.x, d2 = py - q.y, d3 = pz - q.z;
        return d1 * d1 + d2 * d2 + d3 * d3 < q.R;
    } else if (KERNEL == OCT_KERNEL_CIRCLE) {
        double d1 = px - q.x, d2 = py - q.y;
        return d1 * d1 + d2 * d2 <= q.R;
    } else if (KERNEL == OCT_KERNEL_CUBE) {
        return px >= q.lox && px <= q.hix && py >= q.loy && py <= q.hiy && pz >= q.loz && pz <= q.hiz;
    } else {
        return px >= q.lox && px <= q.hix && py >= q.loy && py <= q.hiy;
    }
}

constexpr int kStageBytes = 2048;  // per-wave LDS staging unit of the search kernels (bytes)
constexpr int kUnroll = 4;

}  // namespace oct
