// H6: k nearest neighbours (LinearOctree::knn, linear_octree.hpp:899-942; harness label KNNV2).
//
// Result contract: the min(n, k) points with the smallest squared FP64 distance, the distance
// evaluated exactly as distPointsSquared (linear_octree.hpp:844-851: dx = q.x - p.x, ...,
// dx*dx + dy*dy + dz*dz, no contraction), ordered by (distance, sorted index). The reference's
// heap returns the same distances; only the order inside a tie class is unspecified there.
//
// One wave serves G = 64 / S queries that are consecutive in SFC order (S lanes per query) and
// walks the tree once for all of them, depth first, children pushed so that the octant of the
// wave's middle query comes off the stack first. A child is skipped for a query when a lower
// bound of its box distance exceeds the query's current k-th distance; the bound widens every box
// by 2^-40 of the cloud's magnitude, which covers the FP64 rounding of node centres and of points
// on cell faces, and compares with a relative margin, so pruning never drops a true neighbour.
// Leaf points are staged in LDS once per wave. Each query keeps an LDS buffer of 2 KC entries that
// starts with its best entries so far, sorted, and appends every point that beats its current k-th
// entry. When a buffer could overflow, the wave selects (S lanes per query) and keeps each query's
// k best, which sets the new k-th distance. By merge: the sorted entries own the first KC slots and
// points are appended to the second KC; a selection sorts the second half descending (bitonic) and
// merges it with the first (one bitonic merge of 2 KC). By full sort: points are appended right
// after the kept entries and a selection sorts the whole buffer, as many network stages as a merge
// for KC = 16 but less often when every k of the wave is below KC. Merge serves every KC >= 32 and
// the KC = 16 waves whose largest k is 16 (4 M queries, 8 lanes each, by full sorts vs by merge:
// k = 13 17.7 vs 20.1 ms, k = 14 19.1 vs 20.7, k = 15 20.9 vs 21.2, k = 16 24.0 vs 21.8; one stage per
// round trip, k = 32: 31.3 vs 23.5). The networks run two stages per LDS round
// trip (k = 10: 16.4 -> 15.4 ms FP32; k = 250, 64 lanes: 28.8 -> 24.8 ms for 0.5 M queries), three
// for KC >= 512, where LDS already bounds the occupancy (k = 1000: 37.6 -> 35.0 ms for 0.1 M queries;
// for smaller KC three stages cost a wave per SIMD and run slower).
// FP32_RECHECK filters the points with the leaf-recentred FP32 form
// e = w - 2 a.b + |b|^2 (the count kernel's representation), whose error is below
// hw / 2 = 2^-19 (amax + |b|^2); only points with e <= k-th + hw get their exact FP64 distance.
#include <cfloat>

#include "device.hpp"
#include "search_common.hpp"

namespace oct {

constexpr int kKnnStack = 8 * 22 + 8;  // depth first: at most 7 entries per level + 8 live
constexpr int kKnnChunk = 128;         // leaf points staged per round
constexpr int kKnnMergeMinKC = 32;     // buffer size classes that always select by sort-and-merge
constexpr uint32_t kKnnMergeMinK = 16; // smaller classes: by merge when the wave's largest k reaches this
constexpr int kKnnTripleMinKC = 512;   // buffer size classes whose networks run 3 stages per LDS round trip

struct KnnArgs {
    const double4 *__restrict__ p64;
    const float4 *__restrict__ p32;
    const float *__restrict__ amax;
    const NodeRec *__restrict__ nodes;
    const NodeExt *__restrict__ next;
    const uint32_t *__restrict__ child;
    const uint32_t *__restrict__ order;
    const uint32_t *__restrict__ qidx;
    const double *__restrict__ qxyz;
    uint32_t nq, k;
    uint32_t qoff;  // all-points queries: slot s is sorted point qoff + s
    int32_t *__restrict__ idx_out;
    double *__restrict__ d2_out;
    float *__restrict__ d2f_out;
    unsigned long long *__restrict__ stats;
    uint32_t *__restrict__ err;  // device error word: kErrStack when the traversal stack was full
    // tree geometry per cloud: level radii radtab[cloud * 66 + 3 * level + axis] and the box magnitude
    // magtab[cloud * 3 + axis] = |centre| + radius; for a forest of clouds, tile_info[wave tile] =
    // cloud | valid slots << 20 and roots[cloud] (null: one cloud, root node 0)
    const double *__restrict__ radtab;
    const double *__restrict__ magtab;
    const uint32_t *__restrict__ tile_info;
    const uint32_t *__restrict__ roots;
    const uint32_t *__restrict__ kq;  // per-query k (null: k for every query); k is the row width
    const uint32_t *__restrict__ cloud_off;  // set: indices written local to the tile's cloud
    // work-queue form: device-resident wave-tile count and the zeroed counter waves take tiles from
    const uint32_t *__restrict__ queue_tiles;
    uint32_t *__restrict__ queue_next;
};

__device__ inline bool knn_less(double da, uint32_t ia, double db, uint32_t ib) { return da < db || (da == db && ia < ib); }

// Lower bound of the squared distance from q to any point of the box c +- r, each axis widened by
// dl; the caller compares it against tau (1 + 2^-40), which covers the rounding of this sum.
__device__ inline double gap2(double q, double c, double r, double dl) {
    const double gap = fabs(q - c) - r - dl;
    return gap > 0.0 ? gap * gap : 0.0;
}

// One wave tile (64 / S queries) of a kNN search, run by one wave (a 64-thread block).
template <int S, int KC, int PREC>
__device__ __attribute__((always_inline)) inline void knn_tile(const KnnArgs &a, const uint32_t tile) {
    constexpr int G = 64 / S;                           // queries per wave
    constexpr int CAP = 2 * KC;                         // buffer entries per query
    constexpr int U = S >= 16 ? 1 : 16 / S;             // points per lane per step
    constexpr int STEP = S * U;                         // points per query per step
    constexpr bool F32 = PREC == OCT_PREC_FP32_RECHECK;
    constexpr int STG = kKnnChunk + STEP;               // staged slots incl. the last step's tail
    static_assert(STEP <= KC, "a step must fit after a selection");
    __shared__ double s_bd[G * CAP];
    __shared__ uint32_t s_bi[G * CAP];
    __shared__ uint32_t s_node[kKnnStack];
    __shared__ uint8_t s_lev[kKnnStack];
    __shared__ __attribute__((aligned(16))) unsigned char s_stage[STG * (F32 ? 16 : 32)];
    float4 *s32 = reinterpret_cast<float4 *>(s_stage);
    double4 *s64 = reinterpret_cast<double4 *>(s_stage);

    const uint32_t lane = threadIdx.x;
    const uint32_t qs = lane / S, g = lane % S;
    const uint32_t slot0 = tile * G;
    const uint32_t slot = slot0 + qs;
    uint32_t cloud = 0, nvt = G, root = 0;
    if (a.tile_info) {
        const uint32_t ti = uniform_u32(a.tile_info[tile]);
        cloud = ti & 0xfffffu, nvt = ti >> 20;
        root = uniform_u32(a.roots[cloud]);
    }
    ConstDouble *radp = const_doubles(a.radtab) + 66 * cloud;  // level radii: radp[3 * level + axis]
    ConstDouble *magp = const_doubles(a.magtab) + 3 * cloud;
    const bool valid = slot < a.nq && qs < nvt;
    const uint64_t segmask = (S == 64 ? ~0ull : ((1ull << S) - 1ull)) << (qs * S);
    const uint32_t kw = a.k;  // output row width
    double *bd = s_bd + qs * CAP;
    uint32_t *bi = s_bi + qs * CAP;

    auto wsync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    uint32_t qid = 0, k = kw;
    double qx = 0.0, qy = 0.0, qz = 0.0;
    if (valid) {
        qid = a.order ? a.order[slot] : slot;
        if (a.kq) k = a.kq[qid];
        if (a.qxyz) {
            qx = a.qxyz[3 * (size_t)qid], qy = a.qxyz[3 * (size_t)qid + 1], qz = a.qxyz[3 * (size_t)qid + 2];
        } else {
            const double4 p = a.p64[a.qidx ? a.qidx[qid] : a.qoff + qid];
            qx = p.x, qy = p.y, qz = p.z;
        }
    }
    // box widening per axis: 2^-40 of the magnitudes involved (cloud box and query)
    const double dlx = 0x1p-40 * (magp[0] + fabs(qx)), dly = 0x1p-40 * (magp[1] + fabs(qy));
    const double dlz = 0x1p-40 * (magp[2] + fabs(qz));

    // select by sort-and-merge (else by full sorts); the same in every lane (a ballot)
    const bool merge = KC >= kKnnMergeMinKC || ballot64(valid && k >= kKnnMergeMinK) != 0;
    for (uint32_t t = g; t < (uint32_t)CAP; t += S) bd[t] = __builtin_huge_val(), bi[t] = 0xffffffffu;
    uint32_t cnt = 0;        // kept entries of this query, sorted, at [0, cnt) (same in its S lanes)
    uint32_t fresh = 0;      // appended entries, unsorted, at [KC or cnt, + fresh) (merge or full sort)
    bool full = false;       // k entries seen: tau is the k-th
    double tau = __builtin_huge_val(), tau_pr = __builtin_huge_val();
    uint32_t tau_i = 0xffffffffu;
    float tau32 = __builtin_huge_valf();  // tau rounded up to FP32
    uint64_t n_nodes = 0, n_test = 0, n_recheck = 0, n_select = 0, n_steps = 0;

    // compare-exchange in registers: afterwards (da, ia) <= (db, ib) when up, >= when not
    auto cx = [](double &da, uint32_t &ia, double &db, uint32_t &ib, const bool up) {
        const bool sw = up ? knn_less(db, ib, da, ia) : knn_less(da, ia, db, ib);
        const double lo = sw ? db : da, hi = sw ? da : db;
        const uint32_t li = sw ? ib : ia, hj = sw ? ia : ib;
        da = lo, db = hi, ia = li, ib = hj;
    };
    // bitonic stages with strides `stride`, stride / 2, ..., 1 over the n entries at (sd, si), in
    // blocks of `size` entries alternately ascending and descending (descending first when desc).
    // Two stages per LDS round trip: a lane loads the 4 entries that strides s and s / 2 connect
    // (b, b + s / 2, b + s, b + 3 s / 2), applies both stages in registers and stores them back.
    auto stages = [&](double *sd, uint32_t *si, const uint32_t n, const uint32_t size, uint32_t stride,
                      const bool desc) {
        if constexpr (KC >= kKnnTripleMinKC) {
            // three stages per round trip: the 8 entries that strides s, s / 2 and s / 4 connect
#pragma unroll 1
            for (; stride >= 4; stride >>= 3) {
                const uint32_t h = stride >> 1, q = stride >> 2;
#pragma unroll 1
                for (uint32_t t = g; t < n / 8; t += S) {
                    const uint32_t e0 = ((t & ~(q - 1)) << 3) | (t & (q - 1));
                    const bool up = ((e0 & size) == 0) != desc;
                    const uint32_t o[8] = {0, q, h, h + q, stride, stride + q, stride + h, stride + h + q};
                    double d[8];
                    uint32_t ix[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) d[u] = sd[e0 + o[u]], ix[u] = si[e0 + o[u]];
#pragma unroll
                    for (int u = 0; u < 4; u++) cx(d[u], ix[u], d[u + 4], ix[u + 4], up);
#pragma unroll
                    for (int u = 0; u < 8; u += 4) {
                        cx(d[u], ix[u], d[u + 2], ix[u + 2], up);
                        cx(d[u + 1], ix[u + 1], d[u + 3], ix[u + 3], up);
                    }
#pragma unroll
                    for (int u = 0; u < 8; u += 2) cx(d[u], ix[u], d[u + 1], ix[u + 1], up);
#pragma unroll
                    for (int u = 0; u < 8; u++) sd[e0 + o[u]] = d[u], si[e0 + o[u]] = ix[u];
                }
                wsync();
            }
        }
#pragma unroll 1
        for (; stride >= 2; stride >>= 2) {
            const uint32_t h = stride >> 1;
#pragma unroll 1
            for (uint32_t t = g; t < n / 4; t += S) {
                const uint32_t e0 = ((t & ~(h - 1)) << 2) | (t & (h - 1));
                const uint32_t e1 = e0 + h, e2 = e0 + stride, e3 = e2 + h;
                const bool up = ((e0 & size) == 0) != desc;
                double d0 = sd[e0], d1 = sd[e1], d2 = sd[e2], d3 = sd[e3];
                uint32_t i0 = si[e0], i1 = si[e1], i2 = si[e2], i3 = si[e3];
                cx(d0, i0, d2, i2, up);
                cx(d1, i1, d3, i3, up);
                cx(d0, i0, d1, i1, up);
                cx(d2, i2, d3, i3, up);
                sd[e0] = d0, sd[e1] = d1, sd[e2] = d2, sd[e3] = d3;
                si[e0] = i0, si[e1] = i1, si[e2] = i2, si[e3] = i3;
            }
            wsync();
        }
        if (stride == 1) {  // an odd number of stages: the last one alone
#pragma unroll 1
            for (uint32_t t = g; t < n / 2; t += S) {
                const uint32_t e0 = t << 1;
                const bool up = ((e0 & size) == 0) != desc;
                double d0 = sd[e0], d1 = sd[e0 + 1];
                uint32_t i0 = si[e0], i1 = si[e0 + 1];
                cx(d0, i0, d1, i1, up);
                sd[e0] = d0, sd[e0 + 1] = d1;
                si[e0] = i0, si[e0 + 1] = i1;
            }
            wsync();
        }
    };
    // bitonic sort of the n entries at (sd, si), ascending or descending by (distance, index)
    auto bitonic = [&](double *sd, uint32_t *si, const uint32_t n, const bool descending) {
#pragma unroll 1
        for (uint32_t size = 2; size <= n; size <<= 1) stages(sd, si, n, size, size >> 1, descending);
    };

    // merge every query's appended entries into its kept ones, keep its k best
    auto select = [&]() {
        n_select++;
        wsync();
        if (merge) {
            // appended half [KC, 2 KC) sorted descending (empty slots hold +inf and end up first)
            bitonic(bd + KC, bi + KC, (uint32_t)KC, true);
            // ascending kept half + descending appended half is bitonic: one merge sorts all 2 KC
            stages(bd, bi, (uint32_t)CAP, 2u * CAP, (uint32_t)KC, false);
        } else {
            bitonic(bd, bi, (uint32_t)CAP, false);  // kept and appended entries, [0, cnt + fresh)
        }
        const uint32_t tot = cnt + fresh;  // entries, now sorted at [0, tot)
        const uint32_t keep = tot < k ? tot : k;
        for (uint32_t t = keep + g; t < tot; t += S) bd[t] = __builtin_huge_val(), bi[t] = 0xffffffffu;
        cnt = keep;
        fresh = 0;
        if (keep == k && k > 0) {
            full = true;
            tau = bd[k - 1];
            tau_i = bi[k - 1];
            tau_pr = tau * (1.0 + 0x1p-40);
            tau32 = __double2float_ru(tau);
        }
        wsync();
    };

    // leaf [first, last) with centre c and FP32 bound amx, for the queries with need set
    auto leaf = [&](bool need, uint32_t first, uint32_t last, double cx, double cy, double cz, float amx) {
        float m2x = 0.f, m2y = 0.f, m2z = 0.f, bb = 0.f, hw = 0.f;
        bool force = false;
        if (F32) {
            const float bx = (float)(qx - cx), by = (float)(qy - cy), bz = (float)(qz - cz);
            bb = bx * bx + by * by + bz * bz;
            const float s = amx + bb;
            force = !(s < 1e30f);
            hw = s * 0x1p-18f;
            m2x = -2.0f * bx, m2y = -2.0f * by, m2z = -2.0f * bz;
        }
        if (need && g == 0) n_nodes++;
        for (uint32_t base = first; base < last; base += kKnnChunk) {
            const uint32_t P = (last - base) < (uint32_t)kKnnChunk ? (last - base) : (uint32_t)kKnnChunk;
            const uint32_t nsteps = (P + STEP - 1) / STEP;
            const uint32_t slots = nsteps * STEP;
            for (uint32_t t = lane; t < slots; t += 64) {
                if (F32)
                    s32[t] = t < P ? a.p32[base + t] : make_float4(0.f, 0.f, 0.f, 0.f);
                else
                    s64[t] = t < P ? a.p64[base + t] : make_double4(0.0, 0.0, 0.0, 0.0);
            }
            wsync();
            for (uint32_t s = 0; s < nsteps; s++) {
                // appended entries start at KC (merge) or right after the kept ones (full sort)
                if (ballot64(need && fresh + STEP > (merge ? (uint32_t)KC : (uint32_t)CAP - cnt))) select();
                n_steps++;
                float thr = __builtin_huge_valf();
                if (F32 && full && !force) thr = ((tau32 + hw) * (1.0f + 0x1p-20f));
                double d[U];
                uint32_t id[U];
                bool c[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t t = s * STEP + u * S + g;
                    id[u] = base + t;
                    const bool real = need && t < P;
                    if (F32) {
                        const float4 p = s32[t];
                        const float e = __builtin_fmaf(p.x, m2x, __builtin_fmaf(p.y, m2y, __builtin_fmaf(p.z, m2z, p.w + bb)));
                        c[u] = real && !(e > thr);
                        d[u] = 0.0;
                    } else {
                        const double4 p = s64[t];
                        const double dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                        d[u] = dx * dx + dy * dy + dz * dz;
                        c[u] = real && (!full || knn_less(d[u], id[u], tau, tau_i));
                    }
                    if (real) n_test++;
                }
                if (F32) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (c[u]) {
                            const double4 p = a.p64[id[u]];
                            const double dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                            d[u] = dx * dx + dy * dy + dz * dz;
                            c[u] = !full || knn_less(d[u], id[u], tau, tau_i);
                            n_recheck++;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint64_t m = ballot64(c[u]) & segmask;
                    if (c[u]) {
                        const uint32_t pos = (merge ? (uint32_t)KC : cnt) + fresh + (uint32_t)__popcll(m & lanemask_lt());
                        bd[pos] = d[u];
                        bi[pos] = id[u];
                    }
                    fresh += (uint32_t)__popcll(m);
                }
            }
            wsync();  // the next chunk overwrites the staging area
        }
    };

    // reference query for the child order: the middle valid query of the wave
    // (a forest's wave tile holds nvt valid slots: the rest are padding whose centre is not a query)
    uint32_t nvalid = a.nq - slot0 < (uint32_t)G ? a.nq - slot0 : (uint32_t)G;
    if (nvt < nvalid) nvalid = nvt;
    const int refl = (int)(S * (nvalid / 2));

    uint32_t sp = 0;
    bool overflow = false;  // wave-uniform: a push found the stack full (cannot happen within 21 levels)
    {
        const uint32_t c0 = uniform_u32(a.child[root]);
        if (c0 == 0) {
            const NodeRec nr = a.nodes[root];
            leaf(valid, uniform_u32(nr.first), uniform_u32(nr.last), nr.cx, nr.cy, nr.cz, a.amax[root]);
        } else {
            s_node[0] = (c0 - 1) >> 3, s_lev[0] = 0;
            sp = 1;
        }
    }
    wsync();
    while (sp > 0) {
        sp--;
        const uint32_t ent = uniform_u32(s_node[sp]);
        const uint32_t lv = uniform_u32(s_lev[sp]);
        if (lv & 0x80u) {
            // a leaf: entry = NodeExt index * 8 + combo; re-test it against the current k-th distances
            ConstNodeExt &E = node_ext(a.next, ent >> 3);
            const uint32_t c = ent & 7u, L = lv & 31u;
            const double cx = E.ax[c & 1u], cy = E.ax[2 + ((c >> 1) & 1u)], cz = E.ax[4 + (c >> 2)];
            const double lb = (gap2(qx, cx, radp[3 * L], dlx) + gap2(qy, cy, radp[3 * L + 1], dly)) + gap2(qz, cz, radp[3 * L + 2], dlz);
            const bool need = valid && !(lb > tau_pr);
            if (!ballot64(need)) continue;
            leaf(need, E.first[c], E.last[c], cx, cy, cz, E.amax[c]);
            continue;
        }
        // an internal node at level lv: judge its 8 children (level lv + 1) from per-axis terms
        ConstNodeExt &E = node_ext(a.next, ent);
        const uint32_t d = lv + 1;
        const double rx = radp[3 * d], ry = radp[3 * d + 1], rz = radp[3 * d + 2];
        double gx[2], gy[2], gz[2];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            gx[b] = gap2(qx, E.ax[b], rx, dlx);
            gy[b] = gap2(qy, E.ax[2 + b], ry, dly);
            gz[b] = gap2(qz, E.ax[4 + b], rz, dlz);
        }
        uint32_t needbits = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const double lb = (gx[c & 1] + gy[(c >> 1) & 1]) + gz[c >> 2];
            const bool nd = valid && E.last[c] > E.first[c] && !(lb > tau_pr);
            if (ballot64(nd)) needbits |= 1u << c;
        }
        if (g == 0) n_nodes += 8 * (valid ? 1 : 0);
        needbits = uniform_u32(needbits);  // built from ballots: the same in every lane
        if (needbits == 0) continue;
        // push the farthest octants (relative to the reference query) first
        const uint32_t oc = (qx >= 0.5 * (E.ax[0] + E.ax[1]) ? 1u : 0u) | (qy >= 0.5 * (E.ax[2] + E.ax[3]) ? 2u : 0u) |
                            (qz >= 0.5 * (E.ax[4] + E.ax[5]) ? 4u : 0u);
        const uint32_t qc = readlane_u32(oc, refl);
        constexpr uint32_t kOrder = 0x02416537u;  // nibbles, low first: 7 3 5 6 1 4 2 0
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t c = ((kOrder >> (4 * t)) & 7u) ^ qc;
            if ((needbits >> c) & 1u) {
                if (sp < (uint32_t)kKnnStack) {
                    const uint32_t gc = E.gc[c];
                    s_node[sp] = gc ? (gc - 1) >> 3 : ent * 8 + c;
                    s_lev[sp] = (uint8_t)(gc ? d : (d | 0x80u));
                    sp++;
                } else {
                    overflow = true;
                }
            }
        }
        wsync();
    }
    select();
    if (overflow && lane == 0) atomicOr(a.err, kErrStack);

    if (valid) {
        const uint32_t off = a.cloud_off ? a.cloud_off[cloud] : 0u;  // the tile's cloud holds every neighbour
        for (uint32_t t = g; t < kw; t += S) {
            const bool have = t < cnt;
            a.idx_out[(size_t)qid * kw + t] = have ? (int32_t)(bi[t] - off) : -1;
            if (a.d2_out) a.d2_out[(size_t)qid * kw + t] = have ? bd[t] : __builtin_nan("");
            if (a.d2f_out) a.d2f_out[(size_t)qid * kw + t] = have ? (float)bd[t] : __builtin_nanf("");
        }
    }
    if (a.stats) {
        uint64_t found = (valid && g == 0) ? cnt : 0;
        for (int off = 32; off > 0; off >>= 1) {
            n_nodes += __shfl_xor(n_nodes, off);
            n_test += __shfl_xor(n_test, off);
            n_recheck += __shfl_xor(n_recheck, off);
            found += __shfl_xor(found, off);
        }
        if (lane == 0) {
            atomicAdd(&a.stats[0], (unsigned long long)n_nodes);
            atomicAdd(&a.stats[1], (unsigned long long)n_test);
            atomicAdd(&a.stats[2], (unsigned long long)n_recheck);
            atomicAdd(&a.stats[3], (unsigned long long)found);
            atomicAdd(&a.stats[4], (unsigned long long)n_select);  // wave-wide selections
            atomicAdd(&a.stats[5], (unsigned long long)n_steps);   // wave-wide leaf steps
        }
    }
}

// Occupancy target (waves per SIMD): 6 for the 8-lanes, k <= 16, FP32 shape (80 VGPRs and a 20-byte
// spill instead of 92 VGPRs at 5 waves: 75.1 vs 79.9 ms for all 20 M points of the tile at k = 10);
// every other shape, and the work-queue form (which would spill 56 bytes at that target), keep the
// compiler's choice.
template <int S, int KC, int PREC>
constexpr int knn_waves() {
    return (S == 8 && KC == 16 && PREC == OCT_PREC_FP32_RECHECK) ? 6 : 1;
}

template <int S, int KC, int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(knn_waves<S, KC, PREC>()))) void k_knn(const KnnArgs a) {
    knn_tile<S, KC, PREC>(a, blockIdx.x);
}

// Work-queue form (batches): a grid of resident waves; each takes the next wave tile from a device
// counter (queue_next) until the device-resident tile count (queue_tiles) is reached. Every wave
// leaves once the counter has passed the count.
template <int S, int KC, int PREC>
__global__ __launch_bounds__(64) void k_knn_queue(const KnnArgs a) {
    const uint32_t ntiles = uniform_u32(*a.queue_tiles);
    for (;;) {
        uint32_t t = 0;
        if (lane_id() == 0) t = atomicAdd(a.queue_next, 1u);
        t = readlane_u32(t, 0);
        if (t >= ntiles) return;
        knn_tile<S, KC, PREC>(a, t);
    }
}

template <int S, int KC, int PREC>
static unsigned knn_queue_grid() {
    static unsigned cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_knn_queue<S, KC, PREC>, 64, 0);
    cached = (unsigned)((cus > 0 ? cus : 256) * (per > 0 ? per : 1));
    return cached;
}

template <int S, int KC>
static void launch_sk(const KnnArgs &a, int precision, hipStream_t s) {
    if (a.queue_next) {
        if (precision == OCT_PREC_FP32_RECHECK)
            k_knn_queue<S, KC, OCT_PREC_FP32_RECHECK><<<knn_queue_grid<S, KC, OCT_PREC_FP32_RECHECK>(), 64, 0, s>>>(a);
        else
            k_knn_queue<S, KC, OCT_PREC_FP64><<<knn_queue_grid<S, KC, OCT_PREC_FP64>(), 64, 0, s>>>(a);
        return;
    }
    const unsigned nb = (unsigned)((a.nq + (64 / S) - 1) / (64 / S));
    if (precision == OCT_PREC_FP32_RECHECK)
        k_knn<S, KC, OCT_PREC_FP32_RECHECK><<<nb, 64, 0, s>>>(a);
    else
        k_knn<S, KC, OCT_PREC_FP64><<<nb, 64, 0, s>>>(a);
}

// buffer size class of k (entries per query kept between selections)
static int knn_kc(uint32_t k) {
    int kc = 16;
    while ((uint32_t)kc < k) kc <<= 1;
    return kc;
}

// lanes per query offered for each size class, default first
int knn_lane_options(uint32_t k, int out[3]) {
    switch (knn_kc(k)) {
        case 16: out[0] = 4, out[1] = 2, out[2] = 8; return 3;
        case 32: out[0] = 4, out[1] = 2, out[2] = 8; return 3;
        case 64: out[0] = 8, out[1] = 4, out[2] = 16; return 3;
        case 128: out[0] = 16, out[1] = 8, out[2] = 32; return 3;
        case 256: out[0] = 64, out[1] = 32, out[2] = 16; return 3;
        default: out[0] = 64; return 1;
    }
}

static bool launch_knn_shape(const KnnArgs &a, int kc, int precision, int lanes, hipStream_t s) {
#define OCT_KNN_CASE(KCV, SV)                     \
    if (kc == KCV && lanes == SV) {               \
        launch_sk<SV, KCV>(a, precision, s);      \
        return true;                              \
    }
    OCT_KNN_CASE(16, 2) OCT_KNN_CASE(16, 4) OCT_KNN_CASE(16, 8)
    OCT_KNN_CASE(32, 2) OCT_KNN_CASE(32, 4) OCT_KNN_CASE(32, 8)
    OCT_KNN_CASE(64, 4) OCT_KNN_CASE(64, 8) OCT_KNN_CASE(64, 16)
    OCT_KNN_CASE(128, 8) OCT_KNN_CASE(128, 16) OCT_KNN_CASE(128, 32)
    OCT_KNN_CASE(256, 16) OCT_KNN_CASE(256, 32) OCT_KNN_CASE(256, 64)
    OCT_KNN_CASE(512, 64) OCT_KNN_CASE(1024, 64)
#undef OCT_KNN_CASE
    return false;
}

hipError_t launch_knn(oct_ctx *ctx, const KnnLaunch &L, std::string &err) {
    const size_t nq = L.nq;
    const uint32_t k = L.k;
    const int precision = L.precision;
    if (k > 1024) {
        err = "knn: k above 1024 is not supported";
        return hipErrorInvalidValue;
    }
    if (nq > 0xffffffffull) {
        err = "knn: more than 2^32 - 1 queries in one call";
        return hipErrorInvalidValue;
    }
    if (nq == 0 || k == 0) return hipSuccess;
    KnnArgs a;
    a.p64 = ctx->p64.p;
    a.p32 = ctx->p32.p;
    a.amax = ctx->amax.p;
    a.nodes = ctx->nodes.p;
    a.next = ctx->next.p;
    a.child = ctx->child.p;
    a.order = L.order;
    a.qidx = L.qidx;
    a.qxyz = L.qxyz;
    a.nq = (uint32_t)nq;
    a.k = k;
    a.qoff = L.qoff;
    a.idx_out = L.idx;
    a.d2_out = L.d2;
    a.d2f_out = L.d2f;
    a.stats = L.stats;
    a.err = ctx->derr.p;
    a.radtab = ctx->radtab.p;
    a.magtab = ctx->radtab.p + 66;  // the context's magnitudes follow its level radii
    a.tile_info = L.tile_info;
    a.roots = L.roots;
    a.kq = L.kq;
    a.cloud_off = L.cloud_off;
    a.queue_tiles = L.queue_tiles;
    a.queue_next = L.queue_next;
    if (L.radtab) a.radtab = L.radtab, a.magtab = L.magtab;
    const uint32_t kcls = L.kq && L.kclass ? L.kclass : k;  // selection size class of the launch
    int opts[3];
    knn_lane_options(kcls, opts);
    const int lanes = L.lanes > 0 ? L.lanes : opts[0];
    stage_begin(ctx, ST_KNN);
    if (!launch_knn_shape(a, knn_kc(kcls), precision, lanes, ctx->stream)) {
        err = "knn: lanes per query not offered for this k";
        return hipErrorInvalidValue;
    }
    stage_end(ctx, ST_KNN);
    return hipGetLastError();
}

}  // namespace oct
