// H6: k nearest neighbours (LinearOctree::knn, linear_octree.hpp:899-942; harness label KNNV2).
//
// Result contract: the min(n, k) points with the smallest squared FP64 distance, the distance
// evaluated exactly as distPointsSquared (linear_octree.hpp:844-851: dx = q.x - p.x, ...,
// dx*dx + dy*dy + dz*dz, no contraction), ordered by (distance, sorted index). The reference's
// heap returns the same distances; only the order inside a tie class is unspecified there.
//
// One wave64 per query. Depth-first traversal, children visited nearest-first; a node is skipped
// when its box distance exceeds the current k-th distance by a margin (1e-6 m) that covers FP64
// rounding of node boxes and of points sitting on cell faces, so pruning never drops a true
// neighbour. Leaf points are tested 64 at a time; candidates better than the current k-th go to an
// LDS buffer of capacity C >= k + 64, which is bitonic-sorted under (d2, index) and truncated to k when
// it fills up and at the end.
#include <cfloat>

#include "device.hpp"

namespace oct {

constexpr int kKnnStack = 8 * 23;

struct KnnArgs {
    const double4 *__restrict__ p64;
    const NodeRec *__restrict__ nodes;
    const uint32_t *__restrict__ child;
    const uint32_t *__restrict__ order;
    const uint32_t *__restrict__ qidx;
    const double *__restrict__ qxyz;
    uint32_t nq, k;
    int32_t *__restrict__ idx_out;
    double *__restrict__ d2_out;
    unsigned long long *__restrict__ stats;
    double rad[22][3];
};

__device__ inline bool knn_less(double da, uint32_t ia, double db, uint32_t ib) { return da < db || (da == db && ia < ib); }

// squared distance from q to the box c +- r (distPointOctantSquared, linear_octree.hpp:854-878)
__device__ inline double box_dist2(double qx, double qy, double qz, const NodeRec &n, const double *r) {
    double mnx = n.cx - r[0], mxx = n.cx + r[0];
    double mny = n.cy - r[1], mxy = n.cy + r[1];
    double mnz = n.cz - r[2], mxz = n.cz + r[2];
    double cx = fmax(mnx, fmin(qx, mxx)), cy = fmax(mny, fmin(qy, mxy)), cz = fmax(mnz, fmin(qz, mxz));
    double dx = qx - cx, dy = qy - cy, dz = qz - cz;
    return dx * dx + dy * dy + dz * dz;
}

template <int C>
__device__ inline void bitonic_sort(double *bd, uint32_t *bi) {
    const uint32_t lane = lane_id();
    for (uint32_t size = 2; size <= (uint32_t)C; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (uint32_t t = lane; t < (uint32_t)C / 2; t += 64) {
                uint32_t i = 2 * stride * (t / stride) + (t % stride);
                uint32_t j = i + stride;
                bool up = (i & size) == 0;
                double di = bd[i], dj = bd[j];
                uint32_t ii = bi[i], ij = bi[j];
                bool sw = up ? knn_less(dj, ij, di, ii) : knn_less(di, ii, dj, ij);
                if (sw) {
                    bd[i] = dj, bd[j] = di;
                    bi[i] = ij, bi[j] = ii;
                }
            }
        }
    }
    __syncthreads();
}

template <int C>
__global__ __launch_bounds__(64) void k_knn(const KnnArgs a) {
    __shared__ double bd[C];
    __shared__ uint32_t bi[C];
    __shared__ uint32_t s_node[kKnnStack];
    __shared__ double s_md[kKnnStack];
    __shared__ uint8_t s_dep[kKnnStack];

    const uint32_t slot = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (slot >= a.nq) return;
    const uint32_t qid = a.order ? a.order[slot] : slot;
    double qx, qy, qz;
    if (a.qxyz) {
        qx = a.qxyz[3 * (size_t)qid], qy = a.qxyz[3 * (size_t)qid + 1], qz = a.qxyz[3 * (size_t)qid + 2];
    } else {
        double4 p = a.p64[a.qidx ? a.qidx[qid] : qid];
        qx = p.x, qy = p.y, qz = p.z;
    }
    const uint32_t k = a.k;
    uint32_t cnt = 0;                 // entries in the buffer (uniform)
    double tau = DBL_MAX;             // current k-th distance (valid once full)
    uint32_t tau_i = 0xffffffffu;
    bool full = false;
    uint64_t n_test = 0, n_nodes = 0;

    auto merge = [&]() {
        for (uint32_t t = cnt + lane; t < (uint32_t)C; t += 64) bd[t] = DBL_MAX, bi[t] = 0xffffffffu;
        bitonic_sort<C>(bd, bi);
        if (cnt >= k) {
            cnt = k;
            full = true;
            tau = bd[k - 1];
            tau_i = bi[k - 1];
        }
    };
    auto prune_bound = [&]() -> double {  // (sqrt(tau) + 1e-6)^2
        if (!full) return DBL_MAX;
        double s = sqrt(tau) + 1e-6;
        return s * s;
    };

    uint32_t sp = 0;
    s_node[0] = 0, s_md[0] = 0.0, s_dep[0] = 0;
    sp = 1;
    while (sp > 0) {
        __syncthreads();
        sp--;
        const uint32_t node = s_node[sp];
        const double md = s_md[sp];
        const uint32_t dep = s_dep[sp];
        __syncthreads();
        if (md > prune_bound()) continue;
        n_nodes++;
        const uint32_t fc = a.child[node];
        if (fc == 0) {
            const NodeRec nr = a.nodes[node];
            for (uint32_t base = nr.first; base < nr.last; base += 64) {
                const uint32_t i = base + lane;
                const bool v = i < nr.last;
                double d = DBL_MAX;
                if (v) {
                    const double4 p = a.p64[i];
                    double dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                    d = dx * dx + dy * dy + dz * dz;
                }
                const bool cand = v && (!full || knn_less(d, i, tau, tau_i));
                uint64_t m = ballot64(cand);
                uint32_t nc = (uint32_t)__popcll(m);
                n_test += (uint64_t)__popcll(ballot64(v));
                if (nc == 0) continue;
                if (cnt + nc > (uint32_t)C) {
                    merge();
                    const bool cand2 = cand && (!full || knn_less(d, i, tau, tau_i));
                    m = ballot64(cand2);
                    nc = (uint32_t)__popcll(m);
                    if (nc == 0) continue;
                }
                if (m & (1ull << lane)) {
                    uint32_t pos = cnt + (uint32_t)__popcll(m & lanemask_lt());
                    bd[pos] = d;
                    bi[pos] = i;
                }
                cnt += nc;
            }
        } else {
            // children: lane c < 8 computes its child's box distance; push farthest first
            double cd = DBL_MAX;
            if (lane < 8) cd = box_dist2(qx, qy, qz, a.nodes[fc + lane], a.rad[dep + 1]);
            uint32_t rankc = 0;
            for (uint32_t o = 0; o < 8; o++) {
                double od = __shfl(cd, (int)o);
                if (lane < 8 && (od > cd || (od == cd && o > lane))) rankc++;
            }
            // rankc = number of children to push before this one (farther children first)
            __syncthreads();
            if (lane < 8 && sp + rankc < (uint32_t)kKnnStack) {
                s_node[sp + rankc] = fc + lane;
                s_md[sp + rankc] = cd;
                s_dep[sp + rankc] = (uint8_t)(dep + 1);
            }
            sp = (sp + 8 <= (uint32_t)kKnnStack) ? sp + 8 : (uint32_t)kKnnStack;
        }
    }
    __syncthreads();
    merge();
    const uint32_t found = cnt < k ? cnt : k;
    for (uint32_t t = lane; t < k; t += 64) {
        a.idx_out[(size_t)qid * k + t] = t < found ? (int32_t)bi[t] : -1;
        a.d2_out[(size_t)qid * k + t] = t < found ? bd[t] : __builtin_nan("");
    }
    if (a.stats && lane == 0) {
        atomicAdd(&a.stats[0], (unsigned long long)n_nodes);
        atomicAdd(&a.stats[1], (unsigned long long)n_test);
        atomicAdd(&a.stats[3], (unsigned long long)found);
    }
}

template <int C>
static void launch_c(const KnnArgs &a, hipStream_t s) {
    k_knn<C><<<a.nq, 64, 0, s>>>(a);
}

hipError_t launch_knn(oct_ctx *ctx, int precision, const uint32_t *qidx_dev, const double *qxyz_dev, size_t nq, uint32_t k,
                      int32_t *idx_dev, double *d2_dev, unsigned long long *stats_dev, std::string &err) {
    (void)precision;
    if (k > 1024) {
        err = "knn: k above 1024 is not supported";
        return hipErrorInvalidValue;
    }
    if (nq == 0) return hipSuccess;
    KnnArgs a;
    a.p64 = ctx->p64.p;
    a.nodes = ctx->nodes.p;
    a.child = ctx->child.p;
    a.order = ctx->qidx.p;  // query order staged by the C ABI in qidx[0, nq)
    a.qidx = qidx_dev;
    a.qxyz = qxyz_dev;
    a.nq = (uint32_t)nq;
    a.k = k;
    a.idx_out = idx_dev;
    a.d2_out = d2_dev;
    a.stats = stats_dev;
    for (int l = 0; l < 22; l++)
        for (int c = 0; c < 3; c++) a.rad[l][c] = ctx->level_radii[l][c];
    // capacity: after a merge the buffer holds k entries and must still take a full 64-lane batch
    uint32_t need = k + 64;
    stage_begin(ctx, ST_KNN);
    if (need <= 128)
        launch_c<128>(a, ctx->stream);
    else if (need <= 256)
        launch_c<256>(a, ctx->stream);
    else if (need <= 512)
        launch_c<512>(a, ctx->stream);
    else if (need <= 1024)
        launch_c<1024>(a, ctx->stream);
    else
        launch_c<2048>(a, ctx->stream);
    stage_end(ctx, ST_KNN);
    return hipGetLastError();
}

}  // namespace oct
