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
//  - FP32_RECHECK (sphere, circle), list kernel k_radius: leaf-recentred FP32 coordinates a
//    (w = |a|^2 precomputed) and b = fl32(q - c_leaf); M = fl32(R - |b|^2);
//    u = fma(ax, -2bx, fma(ay, -2by, fma(az, -2bz, w - M))) (circle: fma(ax, ax - 2bx,
//    fma(ay, ay - 2by, -M))) approximates |p - q|^2 - R. Every rounding of this evaluation and of
//    the reference's FP64 one is bounded by hw = 2^-20 (amax + |b|^2 + R)(1 + 2^-10) + |M - thr|,
//    amax = max |a|^2 over the leaf (u32 = 2^-24: |error| <= u32 (13.01 amax + 12.01 |b|^2 + 4 R)
//    + |M - thr| + O(u64)), so u < -hw is inside, u > hw outside, and points with |u| <= hw are
//    re-evaluated with the FP64 expression. In count mode a lane re-counts a whole leaf in FP64 if
//    any point fell in the band.
//  - FP32_RECHECK, count kernel k_count: the same u with the band folded in, t = u + hw, from an
//    FP32 set-up and a wider band hw = 2^-18 (amax + |b|^2 + R) that covers the FP32 set-up's
//    roundings as well (its bound: see the leaf's tests in k_count); counts come from sign bits.
//  - A leaf/query pair whose FP32 terms could overflow, or a leaf holding a non-finite FP32 point
//    (then amax = inf), is decided in FP64 throughout.
//  - Cube/square tests are comparisons only and run on the FP64 coordinates in both modes.
#include <type_traits>

#include "search_common.hpp"

namespace oct {

// Per-lane output sink. MODE 0: neighbour counts only; 1: counts + list-entry counts; 2: fill lists.
template <int ALGO, int MODE>
struct Sink {
    uint32_t cnt = 0;   // neighbour points
    uint32_t ecnt = 0;  // list entries (indices, or ranges for Struct)
    uint64_t base = 0;  // fill offset
    uint32_t *entries = nullptr;

    __device__ inline void range(uint32_t first, uint32_t last) {  // INSIDE node
        if (ALGO == OCT_ALGO_STRUCT || ALGO == OCT_ALGO_APPROX) {  // NeighborSet::addRange keeps empty ranges
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
        if (KERNEL == OCT_KERNEL_CUBE || KERNEL == OCT_KERNEL_SQUARE) box_radii(a, valid, qid, q);
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
        // neighborsApprox beyond its precision level: the node's range whole, or nothing
        // (linear_octree.hpp:987-994)
        if (ALGO == OCT_ALGO_APPROX && d > a.approx_level) return a.approx_upper ? J_INSIDE : J_OUTSIDE;
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
        bool force = false;  // FP32 terms could overflow: every point of the leaf is decided in FP64
        if (kFp32) {
            const float bx = (float)(q.x - nr.cx), by = (float)(q.y - nr.cy), bz = (float)(q.z - nr.cz);
            double b2 = (double)bx * (double)bx + (double)by * (double)by;
            if (KERNEL == OCT_KERNEL_SPHERE) b2 += (double)bz * (double)bz;
            const double am = (double)a.amax[c];
            const double thr = q.R - b2;
            M = (float)thr;
            const double delta = 9.5367431640625e-07 * (am + b2 + q.R) * 1.0009765625 + 1e-30;  // 2^-20 (..)(1 + 2^-10)
            hw = __double2float_ru(delta + fabs((double)M - thr));
            force = !(2.0 * am + 3.0 * b2 + 2.0 * q.R < 1e38);
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
                const bool unc = act && (!(mn > hw) || force);
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
                in = !force && u < -hw;
                if (__builtin_expect(act && !in && (u <= hw || force), 0)) {
                    const double4 pd = a.p64[i];
                    in = inside64<KERNEL>(q, pd.x, pd.y, pd.z);
                    n_recheck++;
                }
            } else {
                const double4 p = a.p64[i];
                in = inside64<KERNEL>(q, p.x, p.y, p.z);
            }
            in = in && act;
            if (ALGO == OCT_ALGO_STRUCT || ALGO == OCT_ALGO_APPROX) {
                if (act && !in) {
                    if (rs < i) sink.run(rs, i);
                    rs = i + 1;
                }
            } else if (in) {
                sink.point(i);
            }
        }
        if ((ALGO == OCT_ALGO_STRUCT || ALGO == OCT_ALGO_APPROX) && act && rs < last) sink.run(rs, last);
    };

    // ---- root (singleTraversal, linear_octree.hpp:655-662)
    uint32_t sp = 0;
    bool overflow = false;  // wave-uniform: a push found the stack full (cannot happen within 21 levels)
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
            // range lists keep empty INSIDE ranges (neighbor_set.hpp:84-89). It still counts as
            // classified: the reference's traversal evaluates its criterion.
            n_class += __popcll(m);
            if (!((ALGO == OCT_ALGO_STRUCT || ALGO == OCT_ALGO_APPROX) && MODE != 0) && uniform_u32(nr.first) == uniform_u32(nr.last))
                continue;
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
                } else {
                    overflow = true;
                }
            }
        }
    }

    if (overflow && lane == 0) atomicOr(a.err, kErrStack);
    if (valid) {
        if (MODE != 2) {
            if (a.counts) a.counts[qid] = sink.cnt;
            if (MODE == 1 && a.ecounts) a.ecounts[qid] = sink.ecnt;
            if (MODE == 1 && a.ecounts64) a.ecounts64[qid] = sink.ecnt;
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

// ------------------------------------------------------------------ count-mode kernel
// Per-query neighbour counts only (the harness keeps result.size(), neighbor_benchmarks.hpp:509).
// G queries per wave, S = 64 / G lanes per query: lane = g * G + j serves query j with "group" g.
//  - the 8 children of a popped node are judged together from per-axis terms (judge8 below);
//  - a leaf's points are staged once per wave in LDS and split across the S lanes of each query
//    (lane group g takes points g, g + S, ...); each staged chunk is padded with sentinel points
//    that are never inside and never in the FP32 band, so every lane runs the same 4x-unrolled
//    trip count with immediate LDS offsets; partial counts are summed at the end;
//  - a smaller G shrinks the union of the wave's traversals (fewer masked-off lane slots), a
//    larger one shares each judgement of 8 children among fewer lanes.
// Empty leaves are skipped: they add nothing to a count whatever their class. Counts do not
// depend on the order of the traversal, which is therefore free.

// Occupancy targets (waves per SIMD), measured on the 50 M-point tile for the sphere kernel with the
// Struct / Prune algorithms: 7 for the FP32 vector path (71 VGPRs, no scratch; at 8 the 64-VGPR and
// 72-SGPR budget spills the children's judgement masks and the search runs 3 % slower), 6 for FP64
// (80 VGPRs). The other variants keep the compiler's choice (forcing them spills registers).
template <int KERNEL, int ALGO, int PREC>
constexpr int count_waves() {
    if (ALGO == OCT_ALGO_NEIGHBORS || !(KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CIRCLE)) return 2;
    return PREC == OCT_PREC_FP32_RECHECK ? 7 : 6;
}
// One wave tile (G queries) of a count-mode search, run by one wave.
template <int KERNEL, int ALGO, int PREC, int G, bool STATS>
__device__ __attribute__((always_inline)) inline void count_tile(const SearchArgs &a, const uint32_t tile) {
    constexpr int S = 64 / G;
    constexpr bool kFp32 = PREC == OCT_PREC_FP32_RECHECK && (KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CIRCLE);
    // real points per staging chunk: 64 FP32 or 48 FP64 points (leaves average ~30 points; a small
    // chunk keeps the LDS per block low enough for 7 (FP32) or 6 (FP64) waves per SIMD: 19.0 / 26.0 KB
    // per 4-wave block). Both take two loads per lane; 96 FP32 points (three) ran 4 % slower, 48
    // FP64 points instead of 32 (fewer chunks for the larger leaves) 4 % faster.
    constexpr int CAP = kFp32 ? kStageBytes / 32 : 48;
    constexpr int SMX = S > 16 ? S : 16;                               // largest lane-group count of a leaf
    constexpr int PAD = kUnroll * SMX + SMX;                           // sentinel + row-padding slots after a chunk
    __shared__ uint32_t s_node[kWaves][kStack];
    __shared__ uint64_t s_mask[kWaves][kStack];
    __shared__ uint8_t s_depth[kWaves][kStack];
    __shared__ uint64_t s_leafm[kWaves][8];  // OVERLAP leaf children of the popped node (query masks)
    __shared__ uint8_t s_rmap[kWaves][64];   // remapped leaf: home slot of compacted query r
    __shared__ uint32_t s_rtot[kWaves][64];  // remapped leaf: count of compacted query r
    __shared__ __attribute__((aligned(16))) unsigned char s_stage[kWaves][(CAP + PAD) * (kFp32 ? 16 : 32)];

    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g = lane / G, j = lane % G;
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    if (uniform_u32(tile * G) >= a.nq) return;
    const uint32_t slot = tile * G + j;
    // one cloud (root 0), or the wave tile's cloud of a forest and its number of valid slots
    uint32_t cloud = 0, nvt = G, root = 0;
    if (a.tile_info) {
        const uint32_t ti = uniform_u32(a.tile_info[tile]);
        cloud = ti & 0xfffffu, nvt = ti >> 20;
        root = uniform_u32(a.roots[cloud]);
    }
    ConstDouble *radp = const_doubles(a.radtab) + 66 * cloud;  // level radii: radp[3 * level + axis]
    const bool valid = slot < a.nq && j < nvt;
    const uint64_t qbit = 1ull << j;
    float4 *sp32 = reinterpret_cast<float4 *>(&s_stage[w][0]);
    double4 *sp64 = reinterpret_cast<double4 *>(&s_stage[w][0]);

    Query q;
    uint32_t qid = 0;
    q.x = q.y = q.z = 0.0;
    q.r = 0.0;
    if (valid) {
        qid = a.order ? a.order[slot] : slot;
        if (a.qxyz) {
            q.x = a.qxyz[3 * (size_t)qid], q.y = a.qxyz[3 * (size_t)qid + 1], q.z = a.qxyz[3 * (size_t)qid + 2];
        } else {
            const double4 p = a.p64[a.qidx ? a.qidx[qid] : qid];
            q.x = p.x, q.y = p.y, q.z = p.z;
        }
        q.r = a.qrad ? a.qrad[qid] : a.radius;
    }
    q.R = q.r * q.r;
    q.lox = q.x - q.r, q.loy = q.y - q.r, q.loz = q.z - q.r;
    q.hix = q.x + q.r, q.hiy = q.y + q.r, q.hiz = q.z + q.r;
    if (KERNEL == OCT_KERNEL_CUBE || KERNEL == OCT_KERNEL_SQUARE) box_radii(a, valid, qid, q);
    const float qRf = (float)q.R;

    uint32_t cnt = 0;
    uint64_t n_class = 0, n_test = 0, n_recheck = 0, n_cslots = 0, n_tslots = 0;

    // a value of another lane (ds_bpermute)
    auto bperm_u32 = [](uint32_t src_lane, uint32_t v) -> uint32_t {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src_lane << 2), (int)v);
    };
    auto bperm_f32 = [&](uint32_t src_lane, float v) -> float { return __uint_as_float(bperm_u32(src_lane, __float_as_uint(v))); };
    auto bperm_f64 = [&](uint32_t src_lane, double v) -> double {
        const uint64_t b = (uint64_t)__double_as_longlong(v);
        const uint64_t lo = bperm_u32(src_lane, (uint32_t)b), hi = bperm_u32(src_lane, (uint32_t)(b >> 32));
        return __longlong_as_double((long long)(lo | (hi << 32)));
    };

    auto judge = [&](bool act, const NodeRec &nr, uint32_t d) -> int {
        if (!act) return -1;
        const double rx = radp[3 * d], ry = radp[3 * d + 1], rz = radp[3 * d + 2];
        if (ALGO == OCT_ALGO_NEIGHBORS) return overlap64<KERNEL>(q, nr.cx, nr.cy, nr.cz, rx, ry, rz) ? J_OVERLAP : J_OUTSIDE;
        return classify64<KERNEL>(q, nr.cx, nr.cy, nr.cz, rx, ry, rz);
    };

    // Stage points [base, base + P) of the sorted cloud for 2^lg lane groups: point t goes to group
    // t % 2^lg, row slot t / 2^lg, rows padded with sentinels to iters = a multiple of kUnroll and
    // spaced iters + 1 slots apart when there are several groups (the spare slot turns the rows'
    // 16-byte columns across LDS banks). Group gp then reads slots gp * (iters + (lg > 0)) + k at
    // fixed offsets. All of a chunk's loads are issued before any LDS write (one round trip);
    // loads read a clamped in-range point and the sentinel is selected by value. Returns iters.
    auto stage = [&](uint32_t base, uint32_t P, uint32_t lg) -> uint32_t {
        const uint32_t st = 1u << lg;
        const uint32_t iters = (P + st * kUnroll - 1) / (st * kUnroll) * kUnroll;  // per-lane iterations
        const uint32_t slots = iters * st;
        const uint32_t rs = iters + (lg ? 1u : 0u);
        constexpr int NI = (CAP + kUnroll * SMX + 63) / 64;  // slots <= CAP + kUnroll * SMX - 1
        if (kFp32) {
            // sentinel: t = 1e30 + C (sphere: w = 1e30) or +inf (circle: x = y = 1e30), never
            // inside and above the band
            const float4 sent = KERNEL == OCT_KERNEL_CIRCLE ? make_float4(1e30f, 1e30f, 0.0f, 0.0f)
                                                            : make_float4(0.0f, 0.0f, 0.0f, 1e30f);
            float4 v[NI];
#pragma unroll
            for (int i = 0; i < NI; i++)
                if (64u * i < slots) v[i] = a.p32[base + min(lane + 64u * i, P - 1)];
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const uint32_t t = lane + 64u * i;
                if (64u * i < slots && t < slots) {
                    const bool real = t < P;
                    sp32[(t & (st - 1u)) * rs + (t >> lg)] =
                        make_float4(real ? v[i].x : sent.x, real ? v[i].y : sent.y, real ? v[i].z : sent.z, real ? v[i].w : sent.w);
                }
            }
        } else {
            double4 v[NI];
#pragma unroll
            for (int i = 0; i < NI; i++)
                if (64u * i < slots) v[i] = a.p64[base + min(lane + 64u * i, P - 1)];
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const uint32_t t = lane + 64u * i;
                if (64u * i < slots && t < slots) {
                    const bool real = t < P;
                    sp64[(t & (st - 1u)) * rs + (t >> lg)] =
                        make_double4(real ? v[i].x : 1e300, real ? v[i].y : 1e300, real ? v[i].z : 1e300, 0.0);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        return iters;
    };

    // leaf: the queries in m count the leaf's points. Normally lane group g of each query takes
    // points g, g + S, ... . A leaf wanted by at most half of the wave's queries is remapped: its mc
    // queries are compacted onto groups of Sp = 64 / Gp lanes (Gp = max(4, mc rounded up to a power
    // of 2), so Sp <= 16); lane l serves compacted query r = l % Gp and takes points l / Gp, + Sp, ...;
    // the query's terms are computed in its own lane and fetched with ds_bpermute, and each group's
    // count returns to its query's first lane through LDS. Same tests and band either way.
    auto leaf = [&](uint64_t m, float am, const NodeRec &nr) {
        const uint32_t first = nr.first, last = nr.last;
        if (STATS) n_test += (uint64_t)__popcll(m) * (last - first);
        const uint32_t mc = (uint32_t)__popcll(m);
        const bool rm = a.remap && 2 * mc <= (uint32_t)G;  // wave-uniform
        constexpr uint32_t LS = S == 1 ? 0 : S == 2 ? 1 : S == 4 ? 2 : S == 8 ? 3 : S == 16 ? 4 : S == 32 ? 5 : 6;
        uint32_t lg = 0, lgs = LS, Sp = S, gp = g, src = lane, rank = 0;
        bool ract = (m & qbit) != 0 && valid;
        const bool home = rm && g == 0 && ((m >> j) & 1ull) != 0;
        if (rm) {
            lg = mc <= 4 ? 2u : 32u - (uint32_t)__builtin_clz(mc - 1);
            lgs = 6u - lg;
            Sp = 64u >> lg;
            gp = lane >> lg;
            const uint32_t r = lane & ((1u << lg) - 1u);
            rank = (uint32_t)__popcll(m & (qbit - 1ull));
            if (home) s_rmap[w][rank] = (uint8_t)j;  // slot j's first lane is lane j
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            ract = r < mc;
            src = ract ? (uint32_t)s_rmap[w][r] : lane;
        }
        // a value of the served query: its own lane's, or fetched from the query's lane
        auto qf = [&](float v) -> float { return rm ? bperm_f32(src, v) : v; };
        auto qd = [&](double v) -> double { return rm ? bperm_f64(src, v) : v; };
        auto fetch_query = [&]() {
            Query qr = q;
            if (KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CIRCLE) {
                qr.x = qd(q.x), qr.y = qd(q.y), qr.R = qd(q.R);
                if (KERNEL == OCT_KERNEL_SPHERE) qr.z = qd(q.z);
            } else {
                qr.lox = qd(q.lox), qr.hix = qd(q.hix), qr.loy = qd(q.loy), qr.hiy = qd(q.hiy);
                if (KERNEL == OCT_KERNEL_CUBE) qr.loz = qd(q.loz), qr.hiz = qd(q.hiz);
            }
            return qr;
        };
        uint32_t lc = 0;
        if (kFp32) {
            // t = fma(ax, -2bx, fma(ay, -2by, fma(az, -2bz, w + C))) (circle: fma(ax, ax - 2bx,
            // fma(ay, ay - 2by, C))) with C = fl32(fl32(hw - fl32(R)) + fl32(|b|^2)) approximates
            // u + hw, u = |p - q|^2 - R. Only b = fl32(q - c) is formed in FP64; the rest of the set-up
            // is FP32. Bounding the error for any association of the five terms, with every product
            // and sum rounded to nearest or toward zero (u' = 2^-23), and with the representation
            // errors of a, b, w and the FP32 set-up (unit u = 2^-24):
            //   |t - (u + hw)| <= u (25 amax + 28 |b|^2 + 13 R + 12 hw) < 32 u (amax + |b|^2 + R);
            // the circle form's fl32(ax - 2bx) adds at most u (2 amax + |b|^2). hw = 2^-18 fl32(fl32(amax +
            // fl32|b|^2) + fl32 R) >= 2 * 32 u (amax + |b|^2 + R)(1 - 5u) covers both with a factor-2
            // margin. Hence t <= 0 (sign bit set, -0 included) is inside, t > 2 hw is outside, and
            // 0 <= t <= 2 hw - an unsigned compare of the bit patterns, which also sends +0 there - is
            // re-tested in FP64. Slots past the leaf are sentinels (t ~ 1e30 or +inf, above 2 hw). A
            // leaf/query pair with amax + |b|^2 + R >= 1e30 (or not finite: a non-finite point makes
            // amax = inf) is re-tested in FP64 outright (lim = ~0), so no term can overflow.
            const float bx = (float)(q.x - nr.cx), by = (float)(q.y - nr.cy), bz = (float)(q.z - nr.cz);
            const float b2 = KERNEL == OCT_KERNEL_SPHERE ? __builtin_fmaf(bx, bx, __builtin_fmaf(by, by, bz * bz))
                                                         : __builtin_fmaf(bx, bx, by * by);
            const float sum = (am + b2) + qRf;
            const float hw = 3.814697265625e-06f * sum;  // 2^-18
            const float C = qf((hw - qRf) + b2);
            const uint32_t lim = __float_as_uint(qf(__uint_as_float(sum < 1e30f ? __float_as_uint(2.0f * hw) : 0xffffffffu)));
            const float ux = qf(-2.0f * bx), uy = qf(-2.0f * by);
            const float uz = KERNEL == OCT_KERNEL_SPHERE ? qf(-2.0f * bz) : 0.0f;
            uint32_t c8 = 0, c8b = 0, mnb = 0xffffffffu;
            for (uint32_t base = first; base < last; base += CAP) {
                const uint32_t P = (last - base) < (uint32_t)CAP ? (last - base) : (uint32_t)CAP;
                const uint32_t iters = stage(base, P, lgs);
                const float4 *pl = sp32 + gp * (iters + (lgs ? 1u : 0u));
                for (uint32_t k = 0; k < iters; k += kUnroll, pl += kUnroll) {
                    const float4 p0 = pl[0], p1 = pl[1], p2 = pl[2], p3 = pl[3];
                    float t0, t1, t2, t3;
                    if (KERNEL == OCT_KERNEL_SPHERE) {
                        t0 = __builtin_fmaf(p0.x, ux, __builtin_fmaf(p0.y, uy, __builtin_fmaf(p0.z, uz, p0.w + C)));
                        t1 = __builtin_fmaf(p1.x, ux, __builtin_fmaf(p1.y, uy, __builtin_fmaf(p1.z, uz, p1.w + C)));
                        t2 = __builtin_fmaf(p2.x, ux, __builtin_fmaf(p2.y, uy, __builtin_fmaf(p2.z, uz, p2.w + C)));
                        t3 = __builtin_fmaf(p3.x, ux, __builtin_fmaf(p3.y, uy, __builtin_fmaf(p3.z, uz, p3.w + C)));
                    } else {
                        t0 = __builtin_fmaf(p0.x, p0.x + ux, __builtin_fmaf(p0.y, p0.y + uy, C));
                        t1 = __builtin_fmaf(p1.x, p1.x + ux, __builtin_fmaf(p1.y, p1.y + uy, C));
                        t2 = __builtin_fmaf(p2.x, p2.x + ux, __builtin_fmaf(p2.y, p2.y + uy, C));
                        t3 = __builtin_fmaf(p3.x, p3.x + ux, __builtin_fmaf(p3.y, p3.y + uy, C));
                    }
                    const uint32_t d0 = __float_as_uint(t0), d1 = __float_as_uint(t1);
                    const uint32_t d2 = __float_as_uint(t2), d3 = __float_as_uint(t3);
                    // two accumulators: each v_bcnt adds into its own (one add3 fewer per 4 tests)
                    c8 = __builtin_popcount(__builtin_amdgcn_perm(d0, d1, 0x0c0c0b09u)) + c8;
                    c8b = __builtin_popcount(__builtin_amdgcn_perm(d2, d3, 0x0c0c0b09u)) + c8b;
                    mnb = min(min(mnb, d0), d1);
                    mnb = min(min(mnb, d2), d3);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            lc = (c8 + c8b) >> 3;
            const bool band = ract && mnb <= lim;
            if (ballot64(band)) {  // rare: exact FP64 re-count of this lane's share of the leaf
                const Query qr = fetch_query();
                if (band) {
                    lc = 0;
                    for (uint32_t i = first + gp; i < last; i += Sp) {
                        const double4 pd = a.p64[i];
                        lc += inside64<KERNEL>(qr, pd.x, pd.y, pd.z);
                        if (STATS) n_recheck++;
                    }
                }
            }
        } else {
            const Query qr = fetch_query();
            for (uint32_t base = first; base < last; base += CAP) {
                const uint32_t P = (last - base) < (uint32_t)CAP ? (last - base) : (uint32_t)CAP;
                const uint32_t iters = stage(base, P, lgs);
                const double4 *pl = sp64 + gp * (iters + (lgs ? 1u : 0u));
                uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
                for (uint32_t k = 0; k < iters; k += kUnroll, pl += kUnroll) {
                    const double4 p0 = pl[0], p1 = pl[1], p2 = pl[2], p3 = pl[3];
                    c0 += inside64<KERNEL>(qr, p0.x, p0.y, p0.z) ? 1u : 0u;
                    c1 += inside64<KERNEL>(qr, p1.x, p1.y, p1.z) ? 1u : 0u;
                    c2 += inside64<KERNEL>(qr, p2.x, p2.y, p2.z) ? 1u : 0u;
                    c3 += inside64<KERNEL>(qr, p3.x, p3.y, p3.z) ? 1u : 0u;
                }
                lc += (c0 + c1) + (c2 + c3);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (!ract) lc = 0;
        if (rm) {
            // butterfly over the lanes that served one compacted query (lane % 2^lg equal, lg >= 2)
            if (lg <= 2) lc += lane_xor<4>(lc);
            if (lg <= 3) lc += lane_xor<8>(lc);
            if (lg <= 4) lc += lane_xor<16>(lc);
            lc += lane_xor32(lc);
            if (ract && gp == 0) s_rtot[w][lane] = lc;  // lane = r for the group's first lane
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            lc = home ? s_rtot[w][rank] : 0u;
        }
        cnt += lc;
        if (STATS) n_tslots += 64ull * ((last - first + Sp - 1) / Sp);
    };

    // The 8 children of a popped node are judged from per-axis terms: they take two centre values
    // per axis (NodeExt), so |q - c|, the kernel-box bounds and the squared extents are computed
    // once per axis value and combined per child, in the reference's operation order, e.g. for the
    // sphere (ex*ex + ey*ey) + ez*ez with ex = |q.x - cx| + rx (kernel_3d.hpp:181-223). Each lane
    // judges all 8 children for its own query (the S lanes of a query agree). Every comparison
    // becomes a lane mask at once (ballot), so the per-child decisions are scalar-unit operations
    // on masks, and the results are lane masks: inm[c] (INSIDE) and ovm[c] (OVERLAP), both within
    // actm, the lanes whose query wants the popped node.
    auto judge8 = [&](ConstNodeExt &E, uint32_t d, uint64_t actm, uint64_t (&inm)[8], uint64_t (&ovm)[8]) {
        if (ALGO == OCT_ALGO_APPROX && d > a.approx_level) {
            // neighborsApprox beyond its precision level: every child whole, or none
            // (linear_octree.hpp:987-994)
#pragma unroll
            for (int c = 0; c < 8; c++) inm[c] = a.approx_upper ? actm : 0ull, ovm[c] = 0ull;
            return;
        }
        const double rx = radp[3 * d], ry = radp[3 * d + 1], rz = radp[3 * d + 2];
        constexpr bool is3d = (KERNEL == OCT_KERNEL_SPHERE || KERNEL == OCT_KERNEL_CUBE);
        if (ALGO == OCT_ALGO_NEIGHBORS || KERNEL == OCT_KERNEL_CUBE || KERNEL == OCT_KERNEL_SQUARE) {
            // boxOverlap (kernel_3d.hpp:100-115, kernel_2d.hpp:29-41) and the box kernels'
            // boxIntersect (kernel_3d.hpp:277-302, kernel_2d.hpp:195-218): per-axis comparisons
            uint64_t ox[2], oy[2], oz[2], ix[2], iy[2], iz[2];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const double hx = E.ax[b] + rx, lx = E.ax[b] - rx, hy = E.ax[2 + b] + ry, ly = E.ax[2 + b] - ry;
                const double hz = E.ax[4 + b] + rz, lz = E.ax[4 + b] - rz;
                ox[b] = ballot64(hx < q.lox || lx > q.hix);
                oy[b] = ballot64(hy < q.loy || ly > q.hiy);
                oz[b] = is3d ? ballot64(hz < q.loz || lz > q.hiz) : 0ull;
                ix[b] = ballot64(hx <= q.hix && lx >= q.lox);
                iy[b] = ballot64(hy <= q.hiy && ly >= q.loy);
                iz[b] = is3d ? ballot64(hz <= q.hiz && lz >= q.loz) : ~0ull;
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
                const uint64_t keep = actm & ~(ox[bx] | oy[by] | oz[bz]);
                const uint64_t inside = ALGO != OCT_ALGO_NEIGHBORS ? keep & ix[bx] & iy[by] & iz[bz] : 0ull;
                inm[c] = inside;
                ovm[c] = keep & ~inside;
            }
        } else {
            // sphere / circle boxIntersect (kernel_3d.hpp:181-223, kernel_2d.hpp:107-144)
            double e2x[2], e2y[2], e2z[2], m2x[2], m2y[2], m2z[2];
            uint64_t ox[2], oy[2], oz[2], nx[2], ny[2], nz[2];
            const double mdx = q.r + rx, mdy = q.r + ry, mdz = q.r + rz;
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const double dx = fabs(q.x - E.ax[b]), dy = fabs(q.y - E.ax[2 + b]);
                ox[b] = ballot64(dx > mdx), oy[b] = ballot64(dy > mdy);
                nx[b] = ballot64(dx < rx), ny[b] = ballot64(dy < ry);
                const double ex = dx + rx, ey = dy + ry;
                e2x[b] = ex * ex, e2y[b] = ey * ey;
                const double fx = dmax0(dx - rx), fy = dmax0(dy - ry);
                m2x[b] = fx * fx, m2y[b] = fy * fy;
                if (is3d) {
                    const double dz = fabs(q.z - E.ax[4 + b]);
                    oz[b] = ballot64(dz > mdz), nz[b] = ballot64(dz < rz);
                    const double ez = dz + rz, fz = dmax0(dz - rz);
                    e2z[b] = ez * ez, m2z[b] = fz * fz;
                } else {
                    oz[b] = nz[b] = 0ull, e2z[b] = m2z[b] = 0.0;
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {  // (bx, by) pairs; each partial sum serves both z halves
                const int bx = k & 1, by = k >> 1;
                const double sxy = e2x[bx] + e2y[by], txy = m2x[bx] + m2y[by];
                const uint64_t oxy = ox[bx] | oy[by], nxy = nx[bx] & ny[by], nxoy = nx[bx] | ny[by];
#pragma unroll
                for (int bz = 0; bz < (is3d ? 2 : 1); bz++) {
                    const int c = k | (bz << 2);
                    const uint64_t keep = actm & ~(oxy | oz[bz]);
                    const double smax = is3d ? sxy + e2z[bz] : sxy;
                    const double smin = is3d ? txy + m2z[bz] : txy;
                    const uint64_t inside = keep & ballot64(smax <= q.R);
                    // two of the three axes inside the node's slab: (nx && ny) || (nz && (nx || ny))
                    const uint64_t two = is3d ? (nxy | (nz[bz] & nxoy)) : nxy;
                    inm[c] = inside;
                    ovm[c] = keep & ~inside & (two | ballot64(smin <= q.R));
                    if (!is3d) {  // 2-D kernels: the z half of the children repeats the judgement
                        inm[c | 4] = inm[c];
                        ovm[c | 4] = ovm[c];
                    }
                }
            }
        }
    };

    // root (singleTraversal, linear_octree.hpp:655-662); group 0 owns the root's INSIDE count.
    // Stack entries hold the FIRST CHILD of an OVERLAP internal node; its NodeExt record is one
    // uniform load.
    uint32_t sp = 0;
    bool overflow = false;  // wave-uniform: a push found the stack full (cannot happen within 21 levels)
    {
        const NodeRec nr = a.nodes[root];
        const int jr = judge(valid, nr, 0);
        if (STATS) n_class += __popcll(ballot64(valid) & gmask), n_cslots += 64;
        if (g == 0 && jr == J_INSIDE) cnt += nr.last - nr.first;
        const uint64_t ov = ballot64(jr == J_OVERLAP) & gmask;
        if (ov) {
            const uint32_t c0 = uniform_u32(a.child[root]);
            if (c0 == 0) {
                NodeRec ur;
                ur.cx = nr.cx, ur.cy = nr.cy, ur.cz = nr.cz;
                ur.first = uniform_u32(nr.first), ur.last = uniform_u32(nr.last);
                if (ur.last > ur.first) leaf(ov, a.amax[root], ur);
            } else {
                s_node[w][0] = c0, s_mask[w][0] = ov, s_depth[w][0] = 0;
                sp = 1;
            }
        }
    }
    while (sp > 0) {
        sp--;
        const uint32_t fc = uniform_u32(s_node[w][sp]);
        const uint64_t m = uniform_u64(s_mask[w][sp]);
        const uint32_t d = uniform_u32(s_depth[w][sp]) + 1;  // the children's level
        ConstNodeExt &E = node_ext(a.next, (fc - 1) >> 3);
        // lanes whose query wants the node: bit j of m for lane g * G + j
        const uint64_t actm = G == 64 ? m : ballot64((m & qbit) != 0);
        uint64_t inl[8], ovl[8];
        judge8(E, d, actm, inl, ovl);
        if (STATS) n_class += 8ull * __popcll(m), n_cslots += 64ull * 8;
        // INSIDE children first: group 0's lanes (0 .. G - 1) hold the counts
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (__builtin_amdgcn_inverse_ballot_w64(inl[c] & gmask)) cnt += E.last[c] - E.first[c];
        uint32_t leaves = 0;  // combos of the OVERLAP non-empty leaf children (bit c)
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t f = E.first[c], l = E.last[c];
            const uint64_t ov = ovl[c] & gmask;  // ballots: uniform already
            if (ov) {
                const uint32_t gc = E.gc[c];
                if (gc == 0) {
                    if (l > f) {  // an empty leaf adds nothing to a count
                        if (lane == 0) s_leafm[w][c] = ov;
                        leaves |= 1u << c;
                    }
                } else if (sp < kStack) {
                    s_node[w][sp] = gc, s_mask[w][sp] = ov, s_depth[w][sp] = (uint8_t)d;
                    sp++;
                } else {
                    overflow = true;
                }
            }
        }
        // the leaves, one code copy: their records are scalar loads from E at a uniform index
        leaves = uniform_u32(leaves);  // built from ballots: the same in every lane
        if (leaves) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        while (leaves) {
            const uint32_t c = (uint32_t)__builtin_ctz(leaves);
            leaves &= leaves - 1;
            const uint64_t ov = uniform_u64(s_leafm[w][c]);
            NodeRec ur;
            ur.cx = E.ax[c & 1], ur.cy = E.ax[2 + ((c >> 1) & 1)], ur.cz = E.ax[4 + ((c >> 2) & 1)];
            ur.first = E.first[c], ur.last = E.last[c];
            leaf(ov, E.amax[c], ur);
        }
    }
    // the S lanes of a query hold parts of its count: lanes j, j + G, ... (G >= 16)
    if constexpr (G <= 16) cnt += lane_xor<16>(cnt);
    if constexpr (G <= 32) cnt += lane_xor32(cnt);
    if (overflow && lane == 0) atomicOr(a.err, kErrStack);
    if (valid && g == 0 && a.counts) a.counts[qid] = cnt;
    if (STATS) {
        if (lane == 0) {
            atomicAdd(&a.stats[0], (unsigned long long)n_class);
            atomicAdd(&a.stats[1], (unsigned long long)n_test);
            atomicAdd(&a.stats[4], (unsigned long long)n_cslots);
            atomicAdd(&a.stats[5], (unsigned long long)n_tslots);
        }
        uint64_t r = n_recheck, nb = (valid && g == 0) ? cnt : 0;
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off), nb += __shfl_xor(nb, off);
        if (lane == 0) {
            if (r) atomicAdd(&a.stats[2], (unsigned long long)r);
            atomicAdd(&a.stats[3], (unsigned long long)nb);
        }
    }
}

template <int KERNEL, int ALGO, int PREC, int G, bool STATS>
__global__ __launch_bounds__(kSearchBlock) __attribute__((amdgpu_waves_per_eu(count_waves<KERNEL, ALGO, PREC>()))) void k_count(const SearchArgs a) {
    count_tile<KERNEL, ALGO, PREC, G, STATS>(a, blockIdx.x * kWaves + (threadIdx.x >> 6));
}

// Work-queue form (batches): a grid of resident waves; each wave takes the next wave tile from a
// device counter (queue_next) until the device-resident tile count (queue_tiles) is reached, so the
// launch shape does not depend on how many tiles the queries made. Every wave leaves once the
// counter has passed the count.
template <int KERNEL, int ALGO, int PREC>
__global__ __launch_bounds__(kSearchBlock) __attribute__((amdgpu_waves_per_eu(count_waves<KERNEL, ALGO, PREC>()))) void k_count_queue(const SearchArgs a) {
    const uint32_t ntiles = uniform_u32(*a.queue_tiles);
    for (;;) {
        uint32_t t = 0;
        if (lane_id() == 0) t = atomicAdd(a.queue_next, 1u);
        t = readlane_u32(t, 0);
        if (t >= ntiles) return;
        count_tile<KERNEL, ALGO, PREC, 64, false>(a, t);
    }
}

template <int KERNEL, int ALGO, int PREC, bool STATS>
static void launch_count_s(const SearchArgs &a, int G, hipStream_t s) {
    const size_t tiles = ((size_t)a.nq + G - 1) / G;
    const unsigned nb = (unsigned)((tiles + kWaves - 1) / kWaves);
    if (G == 16)
        k_count<KERNEL, ALGO, PREC, 16, STATS><<<nb, kSearchBlock, 0, s>>>(a);
    else if (G == 32)
        k_count<KERNEL, ALGO, PREC, 32, STATS><<<nb, kSearchBlock, 0, s>>>(a);
    else
        k_count<KERNEL, ALGO, PREC, 64, STATS><<<nb, kSearchBlock, 0, s>>>(a);
}

// resident waves of the work-queue kernel: blocks per CU at its occupancy, times the CUs
template <int KERNEL, int ALGO, int PREC>
static unsigned queue_grid() {
    static unsigned cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_count_queue<KERNEL, ALGO, PREC>, kSearchBlock, 0);
    cached = (unsigned)((cus > 0 ? cus : 256) * (per > 0 ? per : 1));
    return cached;
}

template <int KERNEL, int ALGO, int PREC>
static void launch_count(const SearchArgs &a, int G, hipStream_t s) {
    if constexpr (ALGO != OCT_ALGO_APPROX) {
        if (a.queue_next) {  // work-queue form: 64 queries per wave tile, vector ALUs, no counters
            k_count_queue<KERNEL, ALGO, PREC><<<queue_grid<KERNEL, ALGO, PREC>(), kSearchBlock, 0, s>>>(a);
            return;
        }
    }
    if constexpr (ALGO == OCT_ALGO_APPROX) {  // one shape: 64 queries per wave, vector ALUs
        const unsigned nb = (unsigned)(((size_t)a.nq + 64 * kWaves - 1) / (64 * kWaves));
        if (a.stats)
            k_count<KERNEL, ALGO, PREC, 64, true><<<nb, kSearchBlock, 0, s>>>(a);
        else
            k_count<KERNEL, ALGO, PREC, 64, false><<<nb, kSearchBlock, 0, s>>>(a);
        return;
    }
    if (a.stats)
        launch_count_s<KERNEL, ALGO, PREC, true>(a, G, s);
    else
        launch_count_s<KERNEL, ALGO, PREC, false>(a, G, s);
}

template <int KERNEL, int ALGO, int PREC>
static void launch_k(const SearchArgs &a, int mode, int G, unsigned nb, hipStream_t s) {
    if (mode == 0 && G > 0)
        launch_count<KERNEL, ALGO, PREC>(a, G, s);
    else if (mode == 0)
        k_radius<KERNEL, ALGO, PREC, 0><<<nb, kSearchBlock, 0, s>>>(a);
    else if (mode == 1)
        k_radius<KERNEL, ALGO, PREC, 1><<<nb, kSearchBlock, 0, s>>>(a);
    else
        k_radius<KERNEL, ALGO, PREC, 2><<<nb, kSearchBlock, 0, s>>>(a);
}

template <int KERNEL, int ALGO>
static void launch_p(const SearchArgs &a, int prec, int mode, int G, unsigned nb, hipStream_t s) {
    if (prec == OCT_PREC_FP32_RECHECK)
        launch_k<KERNEL, ALGO, OCT_PREC_FP32_RECHECK>(a, mode, G, nb, s);
    else
        launch_k<KERNEL, ALGO, OCT_PREC_FP64>(a, mode, G, nb, s);
}

template <int KERNEL>
static void launch_a(const SearchArgs &a, int algo, int prec, int mode, int G, unsigned nb, hipStream_t s) {
    if (algo == OCT_ALGO_NEIGHBORS)
        launch_p<KERNEL, OCT_ALGO_NEIGHBORS>(a, prec, mode, G, nb, s);
    else if (algo == OCT_ALGO_PRUNE)
        launch_p<KERNEL, OCT_ALGO_PRUNE>(a, prec, mode, G, nb, s);
    else if (algo == OCT_ALGO_STRUCT)
        launch_p<KERNEL, OCT_ALGO_STRUCT>(a, prec, mode, G, nb, s);
    else
        launch_p<KERNEL, OCT_ALGO_APPROX>(a, prec, mode, G, nb, s);
}

hipError_t launch_radius(oct_ctx *ctx, const RadiusLaunch &L) {
    if (L.nq == 0) return hipSuccess;
    SearchArgs a;
    a.p64 = ctx->p64.p;
    a.p32 = ctx->p32.p;
    a.amax = ctx->amax.p;
    a.nodes = ctx->nodes.p;
    a.next = ctx->next.p;
    a.child = ctx->child.p;
    a.order = L.order;
    a.qidx = L.qidx;
    a.qxyz = L.qxyz;
    a.qrad = L.qrad;
    a.radius = L.radius;
    a.qrad3 = L.qrad3;
    for (int k = 0; k < 3; k++) a.r3[k] = L.radius3[k];
    a.use_r3 = L.use_r3 ? 1 : 0;
    a.remap = L.remap ? 1 : 0;
    a.approx_level = L.approx_level;
    a.approx_upper = L.approx_upper ? 1 : 0;
    a.radtab = L.radtab ? L.radtab : ctx->radtab.p;
    a.tile_info = L.tile_info;
    a.roots = L.roots;
    a.queue_tiles = L.queue_tiles;
    a.queue_next = L.queue_next;
    a.nq = (uint32_t)L.nq;
    a.counts = L.counts;
    a.ecounts = L.ecounts;
    a.ecounts64 = L.ecounts64;
    a.offsets = L.offsets;
    a.entries = L.entries;
    a.stats = L.stats;
    a.err = ctx->derr.p;
    for (int l = 0; l < 22; l++)
        for (int k = 0; k < 3; k++) a.rad[l][k] = ctx->level_radii[l][k];
    const size_t tiles = (L.nq + 63) / 64;
    const unsigned nb = (unsigned)((tiles + kWaves - 1) / kWaves);
    stage_begin(ctx, ST_SEARCH);
    const int G = L.group;
    switch (L.kernel) {
    case OCT_KERNEL_SPHERE: launch_a<OCT_KERNEL_SPHERE>(a, L.algo, L.precision, L.mode, G, nb, ctx->stream); break;
    case OCT_KERNEL_CIRCLE: launch_a<OCT_KERNEL_CIRCLE>(a, L.algo, L.precision, L.mode, G, nb, ctx->stream); break;
    case OCT_KERNEL_CUBE: launch_a<OCT_KERNEL_CUBE>(a, L.algo, L.precision, L.mode, G, nb, ctx->stream); break;
    default: launch_a<OCT_KERNEL_SQUARE>(a, L.algo, L.precision, L.mode, G, nb, ctx->stream); break;
    }
    stage_end(ctx, ST_SEARCH);
    return hipGetLastError();
}

}  // namespace oct
