// H5: radius searches (neighbors / neighborsPrune / neighborsStruct x sphere / circle / cube /
// square) on gfx950.
//
// Work decomposition: one wave64 = 64 consecutive query slots (SFC-ordered queries are spatially
// coherent). The wave walks the union of its lanes' traversals with one wave-uniform DFS stack in
// LDS; each stack entry carries the mask of lanes for which the node is OVERLAP. Per child, every
// lane in the mask classifies the node for its own query with the reference's FP64 expressions, so
// each lane makes exactly the reference's decisions, and - because the union DFS restricted to one
// lane is that lane's own DFS (same child order, same LIFO stack discipline as singleTraversal,
// linear_octree.hpp:653-686) - each lane emits its neighbours in the reference's order.
// Node records and leaf points are wave-uniform loads (one fetch serves all 64 lanes).
//
// Per-point tests:
//  - FP64: the reference's isInside expression (kernel_3d.hpp:127-138, 232-242; kernel_2d.hpp:53-63,
//    153-162), evaluated as written (no contraction).
//  - FP32_RECHECK (sphere, circle): leaf-recentred FP32 coordinates a (w = |a|^2 precomputed) and
//    b = fl32(q - c_leaf); M = fl32(R - |b|^2); u = fma(ax, -2bx, fma(ay, -2by, fma(az, -2bz, w - M)))
//    (circle: fma(ax, ax - 2bx, fma(ay, ay - 2by, -M))) approximates |p - q|^2 - R. Every rounding of
//    this evaluation and of the reference's FP64 one is bounded by
//    hw = 2^-20 (amax + |b|^2 + R)(1 + 2^-10) + |M - (R - |b|^2)|, amax = max |a|^2 over the leaf
//    (u32 = 2^-24: |error| <= u32 (13.01 amax + 12.01 |b|^2 + 4 R) + |M - thr| + O(u64)), so
//    u < -hw is inside, u > hw outside, and points with |u| <= hw are re-evaluated with the FP64
//    expression. In count mode a lane re-counts a whole leaf in FP64 if any point fell in the band.
//    Cube/square tests are comparisons only and run on the FP64 coordinates in both modes.
#include "device.hpp"

namespace oct {

constexpr int kSearchBlock = 256;
constexpr int kWaves = kSearchBlock / 64;
constexpr int kStack = 8 * 22 + 8;

enum { J_OUTSIDE = 0, J_OVERLAP = 1, J_INSIDE = 2 };

struct SearchArgs {
    const double4 *__restrict__ p64;
    const float4 *__restrict__ p32;
    const float *__restrict__ amax;
    const NodeRec *__restrict__ nodes;
    const uint32_t *__restrict__ child;
    const uint32_t *__restrict__ order;
    const uint32_t *__restrict__ qidx;
    const double *__restrict__ qxyz;
    const double *__restrict__ qrad;
    double radius;
    uint32_t nq;
    uint32_t *__restrict__ counts;
    uint32_t *__restrict__ ecounts;
    const uint64_t *__restrict__ offsets;
    uint32_t *__restrict__ entries;
    unsigned long long *__restrict__ stats;
    double rad[22][3];
};

struct Query {
    double x, y, z, r, R;
    double lox, loy, loz, hix, hiy, hiz;  // kernel box: center -/+ radius (kernel_abstract.hpp:41-45)
};

__device__ inline double dmax0(double a) { return (a < 0.0) ? 0.0 : a; }  // std::max(a, 0.0)

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

// Per-lane output sink. MODE 0: neighbour counts only; 1: counts + list-entry counts; 2: fill lists.
template <int ALGO, int MODE>
struct Sink {
    uint32_t cnt = 0;   // neighbour points
    uint32_t ecnt = 0;  // list entries (indices, or ranges for Struct)
    uint64_t base = 0;  // fill offset
    uint32_t *entries = nullptr;

    __device__ inline void range(uint32_t first, uint32_t last) {  // INSIDE node
        if (ALGO == OCT_ALGO_STRUCT) {                             // NeighborSet::addRange keeps empty ranges
            if (MODE == 2) {
                entries[2 * (base + ecnt)] = first;
                entries[2 * (base + ecnt) + 1] = last;
            }
            ecnt++;
            cnt += last - first;
        } else {
            if (MODE == 2)
                for (uint32_t i = first; i < last; i++) entries[base + ecnt + (i - first)] = i;
            ecnt += last - first;
            cnt += last - first;
        }
    }
    __device__ inline void point(uint32_t i) {  // Prune / neighbors leaf hit
        if (MODE == 2) entries[base + ecnt] = i;
        ecnt++;
        cnt++;
    }
    __device__ inline void run(uint32_t first, uint32_t last) {  // Struct leaf run (non-empty)
        if (MODE == 2) {
            entries[2 * (base + ecnt)] = first;
            entries[2 * (base + ecnt) + 1] = last;
        }
        ecnt++;
        cnt += last - first;
    }
};

template <int KERNEL, int ALGO, int PREC, int MODE>
__global__ __launch_bounds__(kSearchBlock) void k_radius(const SearchArgs a) {
    __shared__ uint32_t s_node[kWaves][kStack];
    __shared__ uint64_t s_mask[kWaves][kStack];
    __shared__ uint8_t s_depth[kWaves][kStack];

    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lanebit = 1ull << lane;
    const uint32_t tile = blockIdx.x * kWaves + w;
    const uint32_t slot = tile * 64 + lane;
    const bool valid = slot < a.nq;
    if (uniform_u32(tile * 64) >= a.nq) return;

    // ---- per-lane query
    Query q;
    uint32_t qid = 0;
    {
        q.x = q.y = q.z = 0.0;
        q.r = 0.0;
        if (valid) {
            qid = a.order ? a.order[slot] : slot;
            if (a.qxyz) {
                q.x = a.qxyz[3 * (size_t)qid], q.y = a.qxyz[3 * (size_t)qid + 1], q.z = a.qxyz[3 * (size_t)qid + 2];
            } else {
                uint32_t pi = a.qidx ? a.qidx[qid] : qid;
                double4 p = a.p64[pi];
                q.x = p.x, q.y = p.y, q.z = p.z;
            }
            q.r = a.qrad ? a.qrad[qid] : a.radius;
        }
        q.R = q.r * q.r;
        q.lox = q.x - q.r, q.loy = q.y - q.r, q.loz = q.z - q.r;
        q.hix = q.x + q.r, q.hiy = q.y + q.r, q.hiz = q.z + q.r;
    }
    Sink<ALGO, MODE> sink;
    if (MODE == 2 && valid) {
        sink.base = a.offsets[qid];
        sink.entries = a.entries;
    }
    uint64_t n_class = 0, n_test = 0, n_recheck = 0, n_cslots = 0, n_tslots = 0;

    // ---- classification helper: returns the lane's judgement for node c at depth d
    auto judge = [&](uint64_t m, const NodeRec &nr, uint32_t d) -> int {
        if (!(m & lanebit)) return -1;
        if (ALGO == OCT_ALGO_NEIGHBORS)
            return overlap64<KERNEL>(q, nr.cx, nr.cy, nr.cz, a.rad[d][0], a.rad[d][1], a.rad[d][2]) ? J_OVERLAP : J_OUTSIDE;
        return classify64<KERNEL>(q, nr.cx, nr.cy, nr.cz, a.rad[d][0], a.rad[d][1], a.rad[d][2]);
    };

    // ---- leaf: lanes in m test the leaf's points (endpointAction, linear_octree.hpp:716-732, 771-780, 804-813)
    auto leaf = [&](uint64_t m, uint32_t c, const NodeRec &nr) {
        const bool act = (m & lanebit) != 0;
        const uint32_t first = uniform_u32(nr.first), last = uniform_u32(nr.last);
        n_test += (uint64_t)__popcll(m) * (last - first);
        n_tslots += 64ull * (last - first);
        constexpr bool kFp32 = PREC == OCT_PREC_FP32_RECHECK && (KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CIRCLE);
        // FP32 band setup (per lane, per leaf). u = w - 2 a.b - M approximates |p - q|^2 - R with
        // |error| <= hw (see the file header), so u < -hw is inside, u > hw outside, else recheck.
        float M = 0.f, hw = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
        if (kFp32) {
            const float bx = (float)(q.x - nr.cx), by = (float)(q.y - nr.cy), bz = (float)(q.z - nr.cz);
            double b2 = (double)bx * (double)bx + (double)by * (double)by;
            if (KERNEL == OCT_KERNEL_SPHERE) b2 += (double)bz * (double)bz;
            const double am = (double)a.amax[c];
            const double thr = q.R - b2;
            M = (float)thr;
            const double delta = 9.5367431640625e-07 * (am + b2 + q.R) * 1.0009765625 + 1e-30;  // 2^-20 (..)(1 + 2^-10)
            hw = __double2float_ru(delta + fabs((double)M - thr));
            ux = -2.0f * bx, uy = -2.0f * by, uz = -2.0f * bz;
        }
        auto u32f = [&](const float4 &p) -> float {
            if (KERNEL == OCT_KERNEL_SPHERE) return __builtin_fmaf(p.x, ux, __builtin_fmaf(p.y, uy, __builtin_fmaf(p.z, uz, p.w - M)));
            return __builtin_fmaf(p.x, p.x + ux, __builtin_fmaf(p.y, p.y + uy, -M));
        };
        if (MODE == 0) {
            // count only: no per-point control flow. cin counts t < L; mn tracks min |t - mid| over
            // the leaf, so a lane has a point inside the band [L, H] iff mn <= hw.
            uint32_t cin = 0;
            if (kFp32) {
                const float nhw = -hw;
                float mn = __builtin_huge_valf();
                uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;  // one counter per unroll slot: compare + carry-add
                uint32_t i = first;
                for (; i + 4 <= last; i += 4) {
                    const float4 p0 = a.p32[i], p1 = a.p32[i + 1], p2 = a.p32[i + 2], p3 = a.p32[i + 3];
                    const float u0 = u32f(p0), u1 = u32f(p1), u2 = u32f(p2), u3 = u32f(p3);
                    c0 += (u0 < nhw) ? 1u : 0u;
                    c1 += (u1 < nhw) ? 1u : 0u;
                    c2 += (u2 < nhw) ? 1u : 0u;
                    c3 += (u3 < nhw) ? 1u : 0u;
                    mn = fminf(mn, fminf(fabsf(u0), fabsf(u1)));
                    mn = fminf(mn, fminf(fabsf(u2), fabsf(u3)));
                }
                for (; i < last; i++) {
                    const float u = u32f(a.p32[i]);
                    c0 += (u < nhw) ? 1u : 0u;
                    mn = fminf(mn, fabsf(u));
                }
                cin = (c0 + c1) + (c2 + c3);
                const bool unc = act && !(mn > hw);
                if (ballot64(unc)) {  // rare: re-count this leaf exactly for the lanes with a band hit
                    if (unc) {
                        n_recheck += last - first;
                        cin = 0;
                        for (uint32_t j = first; j < last; j++) {
                            const double4 pd = a.p64[j];
                            cin += inside64<KERNEL>(q, pd.x, pd.y, pd.z);
                        }
                    }
                }
            } else {
                uint32_t i = first;
                for (; i + 4 <= last; i += 4) {
                    const double4 p0 = a.p64[i], p1 = a.p64[i + 1], p2 = a.p64[i + 2], p3 = a.p64[i + 3];
                    cin += inside64<KERNEL>(q, p0.x, p0.y, p0.z);
                    cin += inside64<KERNEL>(q, p1.x, p1.y, p1.z);
                    cin += inside64<KERNEL>(q, p2.x, p2.y, p2.z);
                    cin += inside64<KERNEL>(q, p3.x, p3.y, p3.z);
                }
                for (; i < last; i++) {
                    const double4 p = a.p64[i];
                    cin += inside64<KERNEL>(q, p.x, p.y, p.z);
                }
            }
            if (act) sink.cnt += cin;
            return;
        }
        // list modes: per-point decisions in the reference's order
        uint32_t rs = first;  // Struct run start
        for (uint32_t i = first; i < last; i++) {
            bool in;
            if (kFp32) {
                const float u = u32f(a.p32[i]);
                in = u < -hw;
                if (__builtin_expect(act && !in && u <= hw, 0)) {
                    const double4 pd = a.p64[i];
                    in = inside64<KERNEL>(q, pd.x, pd.y, pd.z);
                    n_recheck++;
                }
            } else {
                const double4 p = a.p64[i];
                in = inside64<KERNEL>(q, p.x, p.y, p.z);
            }
            in = in && act;
            if (ALGO == OCT_ALGO_STRUCT) {
                if (act && !in) {
                    if (rs < i) sink.run(rs, i);
                    rs = i + 1;
                }
            } else if (in) {
                sink.point(i);
            }
        }
        if (ALGO == OCT_ALGO_STRUCT && act && rs < last) sink.run(rs, last);
    };

    // ---- root (singleTraversal, linear_octree.hpp:655-662)
    uint32_t sp = 0;
    {
        const NodeRec nr = a.nodes[0];
        const uint64_t m0 = ballot64(valid);
        n_class += __popcll(m0);
        n_cslots += 64;
        const int j = judge(m0, nr, 0);
        if (j == J_INSIDE) sink.range(nr.first, nr.last);
        const uint64_t ov = ballot64(j == J_OVERLAP);
        if (ov) {
            if (a.child[0] == 0) {
                leaf(ov, 0, nr);
            } else {
                s_node[w][0] = 0, s_mask[w][0] = ov, s_depth[w][0] = 0;
                sp = 1;
            }
        }
    }
    // ---- DFS: pop, classify the 8 children, emit INSIDE ranges and OVERLAP leaves in octant order,
    //      push OVERLAP internal children (they are popped last-pushed-first, as the reference does)
    while (sp > 0) {
        sp--;
        const uint32_t node = uniform_u32(s_node[w][sp]);
        const uint64_t m = uniform_u64(s_mask[w][sp]);
        const uint32_t d = uniform_u32(s_depth[w][sp]) + 1;
        const uint32_t first_child = uniform_u32(a.child[node]);
        for (uint32_t oct = 0; oct < 8; oct++) {
            const uint32_t c = first_child + oct;
            const NodeRec nr = a.nodes[c];
            // An empty child cannot change counts or index lists whatever its class; only Struct
            // range lists keep empty INSIDE ranges (neighbor_set.hpp:84-89).
            if (!(ALGO == OCT_ALGO_STRUCT && MODE != 0) && uniform_u32(nr.first) == uniform_u32(nr.last)) continue;
            n_class += __popcll(m);
            n_cslots += 64;
            const int j = judge(m, nr, d);
            if (j == J_INSIDE) sink.range(nr.first, nr.last);
            const uint64_t ov = ballot64(j == J_OVERLAP);
            if (ov) {
                const uint32_t cc = uniform_u32(a.child[c]);
                if (cc == 0) {
                    leaf(ov, c, nr);
                } else if (sp < kStack) {
                    s_node[w][sp] = c, s_mask[w][sp] = ov, s_depth[w][sp] = (uint8_t)d;
                    sp++;
                }
            }
        }
    }

    if (valid) {
        if (MODE != 2) {
            if (a.counts) a.counts[qid] = sink.cnt;
            if (MODE == 1 && a.ecounts) a.ecounts[qid] = sink.ecnt;
        }
    }
    if (a.stats && lane == 0) {
        atomicAdd(&a.stats[0], (unsigned long long)n_class);
        atomicAdd(&a.stats[1], (unsigned long long)n_test);
        atomicAdd(&a.stats[4], (unsigned long long)n_cslots);
        atomicAdd(&a.stats[5], (unsigned long long)n_tslots);
    }
    if (a.stats && PREC == OCT_PREC_FP32_RECHECK) {
        // n_recheck is per lane
        uint64_t s = n_recheck;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0 && s) atomicAdd(&a.stats[2], (unsigned long long)s);
    }
    if (a.stats) {
        uint64_t s = valid ? sink.cnt : 0;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) atomicAdd(&a.stats[3], (unsigned long long)s);
    }
}

template <int KERNEL, int ALGO, int PREC>
static void launch_k(const SearchArgs &a, int mode, unsigned nb, hipStream_t s) {
    if (mode == 0)
        k_radius<KERNEL, ALGO, PREC, 0><<<nb, kSearchBlock, 0, s>>>(a);
    else if (mode == 1)
        k_radius<KERNEL, ALGO, PREC, 1><<<nb, kSearchBlock, 0, s>>>(a);
    else
        k_radius<KERNEL, ALGO, PREC, 2><<<nb, kSearchBlock, 0, s>>>(a);
}

template <int KERNEL, int ALGO>
static void launch_p(const SearchArgs &a, int prec, int mode, unsigned nb, hipStream_t s) {
    if (prec == OCT_PREC_FP32_RECHECK)
        launch_k<KERNEL, ALGO, OCT_PREC_FP32_RECHECK>(a, mode, nb, s);
    else
        launch_k<KERNEL, ALGO, OCT_PREC_FP64>(a, mode, nb, s);
}

template <int KERNEL>
static void launch_a(const SearchArgs &a, int algo, int prec, int mode, unsigned nb, hipStream_t s) {
    if (algo == OCT_ALGO_NEIGHBORS)
        launch_p<KERNEL, OCT_ALGO_NEIGHBORS>(a, prec, mode, nb, s);
    else if (algo == OCT_ALGO_PRUNE)
        launch_p<KERNEL, OCT_ALGO_PRUNE>(a, prec, mode, nb, s);
    else
        launch_p<KERNEL, OCT_ALGO_STRUCT>(a, prec, mode, nb, s);
}

hipError_t launch_radius(oct_ctx *ctx, const RadiusLaunch &L) {
    if (L.nq == 0) return hipSuccess;
    SearchArgs a;
    a.p64 = ctx->p64.p;
    a.p32 = ctx->p32.p;
    a.amax = ctx->amax.p;
    a.nodes = ctx->nodes.p;
    a.child = ctx->child.p;
    a.order = L.order;
    a.qidx = L.qidx;
    a.qxyz = L.qxyz;
    a.qrad = L.qrad;
    a.radius = L.radius;
    a.nq = (uint32_t)L.nq;
    a.counts = L.counts;
    a.ecounts = L.ecounts;
    a.offsets = L.offsets;
    a.entries = L.entries;
    a.stats = L.stats;
    for (int l = 0; l < 22; l++)
        for (int k = 0; k < 3; k++) a.rad[l][k] = ctx->level_radii[l][k];
    const size_t tiles = (L.nq + 63) / 64;
    const unsigned nb = (unsigned)((tiles + kWaves - 1) / kWaves);
    stage_begin(ctx, ST_SEARCH);
    switch (L.kernel) {
    case OCT_KERNEL_SPHERE: launch_a<OCT_KERNEL_SPHERE>(a, L.algo, L.precision, L.mode, nb, ctx->stream); break;
    case OCT_KERNEL_CIRCLE: launch_a<OCT_KERNEL_CIRCLE>(a, L.algo, L.precision, L.mode, nb, ctx->stream); break;
    case OCT_KERNEL_CUBE: launch_a<OCT_KERNEL_CUBE>(a, L.algo, L.precision, L.mode, nb, ctx->stream); break;
    default: launch_a<OCT_KERNEL_SQUARE>(a, L.algo, L.precision, L.mode, nb, ctx->stream); break;
    }
    stage_end(ctx, ST_SEARCH);
    return hipGetLastError();
}

}  // namespace oct
