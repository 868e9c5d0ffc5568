// Batches of independent clouds (no counterpart in the reference, whose LinearOctree holds one cloud;
// the reference harness runs its clouds one after another).
//
// Each cloud is encoded, sorted and built by its own context, exactly as a single cloud (same keys,
// permutation and tree arrays). oct_batch_build then lays the clouds out as one forest in a context
// of its own: sorted points concatenated, every cloud's node arrays copied to a node offset that is
// a multiple of 8 with point ranges and child indices rebased, so the count and kNN kernels walk any
// cloud from its root with the unchanged child -> NodeExt rule ((first child - 1) / 8). A search
// over many clouds is then one launch per search kind: queries are grouped per cloud into wave tiles
// (64 slots for counts, 64 / lanes for kNN), and each tile names its cloud, which selects the root
// and that cloud's level radii.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "device.hpp"

struct oct_batch {
    int device = 0;
    std::string err;
    oct_ctx *forest = nullptr;  // the concatenated clouds; its stream runs every batch search
    struct Cloud {
        oct_ctx *ctx = nullptr;
        int encoder = 0;
        uint32_t max_leaf = 0;
        size_t n = 0, pt_off = 0, node_off = 0;
        oct_tree_info info{};
    };
    std::vector<Cloud> clouds;
    bool built = false;
    std::vector<double> h_rad, h_mag;
    std::vector<uint32_t> h_roots;
    oct::DevBuf<double> radtab, magtab;
    oct::DevBuf<uint32_t> roots, tile_info, order, qidx, kq, counts;
    oct::DevBuf<double> qrad;
    oct::DevBuf<int32_t> kidx;
    oct::DevBuf<double> kd2;
};

namespace oct {

// node records of one cloud, point ranges shifted by pt_off
__global__ void k_rebase_nodes(const NodeRec *__restrict__ src, uint32_t n, uint32_t pt_off, NodeRec *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    NodeRec r = src[i];
    r.first += pt_off, r.last += pt_off;
    dst[i] = r;
}
// first-child indices shifted by node_off (0 stays "leaf")
__global__ void k_rebase_child(const uint32_t *__restrict__ src, uint32_t n, uint32_t node_off, uint32_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = src[i];
    dst[i] = c ? c + node_off : 0u;
}
// NodeExt records: point ranges shifted by pt_off, grandchild indices by node_off
__global__ void k_rebase_ext(const NodeExt *__restrict__ src, uint32_t n, uint32_t pt_off, uint32_t node_off,
                             NodeExt *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    NodeExt e = src[i];
    for (int c = 0; c < 8; c++) {
        e.first[c] += pt_off, e.last[c] += pt_off;
        e.gc[c] = e.gc[c] ? e.gc[c] + node_off : 0u;
    }
    dst[i] = e;
}

}  // namespace oct

using namespace oct;

static int bfail(oct_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}
#define BCHK(call, what)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return bfail(b, OCT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BAPI(call, what)                                                                                   \
    do {                                                                                                   \
        int r_ = (call);                                                                                   \
        if (r_ != OCT_OK) return bfail(b, r_, std::string(what) + ": " + oct_last_error(ctx_for_error)); \
    } while (0)

extern "C" {

int oct_batch_create(oct_batch **out, int device) {
    if (!out) return OCT_ERR_ARG;
    *out = nullptr;
    oct_ctx *f = nullptr;
    int r = oct_create(&f, device);
    if (r != OCT_OK) return r;
    oct_batch *b = new oct_batch;
    b->device = device;
    b->forest = f;
    *out = b;
    return OCT_OK;
}

int oct_batch_destroy(oct_batch *b) {
    if (!b) return OCT_ERR_ARG;
    for (auto &c : b->clouds)
        if (c.ctx) oct_destroy(c.ctx);
    hipSetDevice(b->device);
    hipStreamSynchronize(b->forest->stream);
    b->radtab.release(), b->magtab.release(), b->roots.release(), b->tile_info.release(), b->order.release();
    b->qidx.release(), b->kq.release(), b->counts.release(), b->qrad.release(), b->kidx.release(), b->kd2.release();
    oct_destroy(b->forest);
    delete b;
    return OCT_OK;
}

const char *oct_batch_last_error(const oct_batch *b) { return b ? b->err.c_str() : "null batch"; }

int oct_batch_add(oct_batch *b, const double *xyz, size_t n, size_t stride_bytes, int encoder, uint32_t max_points_leaf,
                  int *cloud_id) {
    if (!b) return OCT_ERR_ARG;
    if (b->built) return bfail(b, OCT_ERR_STATE, "batch already built");
    if (encoder != OCT_ENC_MORTON3D && encoder != OCT_ENC_HILBERT3D) return bfail(b, OCT_ERR_UNSUPPORTED, "encoder");
    if (max_points_leaf == 0) return bfail(b, OCT_ERR_ARG, "max_points_leaf must be positive");
    if (b->clouds.size() >= (1u << 20) - 1) return bfail(b, OCT_ERR_ARG, "too many clouds in one batch");
    oct_batch::Cloud c;
    int r = oct_create(&c.ctx, b->device);
    if (r != OCT_OK) return bfail(b, r, "context of a cloud");
    oct_ctx *ctx_for_error = c.ctx;
    if ((r = oct_set_points(c.ctx, xyz, n, stride_bytes)) != OCT_OK) {
        std::string m = oct_last_error(ctx_for_error);
        oct_destroy(c.ctx);
        return bfail(b, r, "set points: " + m);
    }
    c.encoder = encoder, c.max_leaf = max_points_leaf, c.n = n;
    b->clouds.push_back(c);
    if (cloud_id) *cloud_id = (int)b->clouds.size() - 1;
    return OCT_OK;
}

int oct_batch_build(oct_batch *b) {
    if (!b) return OCT_ERR_ARG;
    if (b->built) return bfail(b, OCT_ERR_STATE, "batch already built");
    BCHK(hipSetDevice(b->device), "set device");
    // every cloud on its own: bbox, encode, sort, build (the single-cloud path, bit-identical)
    size_t pts = 0, nodes = 0;
    for (auto &c : b->clouds) {
        oct_ctx *ctx_for_error = c.ctx;
        double cen[3], rad[3];
        BAPI(oct_sort_points(c.ctx, c.encoder, cen, rad, nullptr, nullptr), "sort a cloud");
        BAPI(oct_build(c.ctx, c.max_leaf, &c.info), "build a cloud");
        BCHK(hipStreamSynchronize(c.ctx->stream), "build a cloud");
        c.pt_off = pts, c.node_off = nodes;
        pts += c.n;
        nodes += ((size_t)c.info.n_total + 7) / 8 * 8;
    }
    if (pts >= 0xffffffffull || nodes >= 0xffffffffull) return bfail(b, OCT_ERR_ARG, "batch beyond 2^32 points or nodes");
    // the forest
    oct_ctx *f = b->forest;
    BCHK(f->p64.reserve(pts + 1), "alloc forest points");
    BCHK(f->p32.reserve(pts + 1), "alloc forest points");
    BCHK(f->nodes.reserve(nodes + 8), "alloc forest nodes");
    BCHK(f->child.reserve(nodes + 8), "alloc forest nodes");
    BCHK(f->amax.reserve(nodes + 8), "alloc forest nodes");
    BCHK(f->next.reserve(nodes / 8 + 1), "alloc forest nodes");
    const size_t C = b->clouds.size();
    b->h_rad.assign(66 * (C ? C : 1), 0.0), b->h_mag.assign(3 * (C ? C : 1), 0.0), b->h_roots.assign(C ? C : 1, 0u);
    for (size_t i = 0; i < C; i++) {
        const auto &c = b->clouds[i];
        oct_ctx *x = c.ctx;
        const uint32_t nt = c.info.n_total, ni = c.info.n_internal;
        if (c.n) {
            BCHK(hipMemcpyAsync(f->p64.p + c.pt_off, x->p64.p, c.n * sizeof(double4), hipMemcpyDeviceToDevice, f->stream), "copy");
            BCHK(hipMemcpyAsync(f->p32.p + c.pt_off, x->p32.p, c.n * sizeof(float4), hipMemcpyDeviceToDevice, f->stream), "copy");
        }
        BCHK(hipMemcpyAsync(f->amax.p + c.node_off, x->amax.p, nt * sizeof(float), hipMemcpyDeviceToDevice, f->stream), "copy");
        const unsigned nb = (nt + 255) / 256;
        k_rebase_nodes<<<nb, 256, 0, f->stream>>>(x->nodes.p, nt, (uint32_t)c.pt_off, f->nodes.p + c.node_off);
        k_rebase_child<<<nb, 256, 0, f->stream>>>(x->child.p, nt, (uint32_t)c.node_off, f->child.p + c.node_off);
        if (ni)
            k_rebase_ext<<<(ni + 255) / 256, 256, 0, f->stream>>>(x->next.p, ni, (uint32_t)c.pt_off, (uint32_t)c.node_off,
                                                                  f->next.p + c.node_off / 8);
        BCHK(hipGetLastError(), "rebase");
        for (int l = 0; l < 22; l++)
            for (int a = 0; a < 3; a++) b->h_rad[66 * i + 3 * l + a] = x->level_radii[l][a];
        for (int a = 0; a < 3; a++) b->h_mag[3 * i + a] = std::fabs(x->center[a]) + std::fabs(x->radii[a]);
        b->h_roots[i] = (uint32_t)c.node_off;
    }
    BCHK(b->radtab.reserve(b->h_rad.size()), "alloc");
    BCHK(b->magtab.reserve(b->h_mag.size()), "alloc");
    BCHK(b->roots.reserve(b->h_roots.size()), "alloc");
    BCHK(hipMemcpyAsync(b->radtab.p, b->h_rad.data(), b->h_rad.size() * 8, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->magtab.p, b->h_mag.data(), b->h_mag.size() * 8, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->roots.p, b->h_roots.data(), b->h_roots.size() * 4, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipStreamSynchronize(f->stream), "assemble the forest");
    f->n = pts;
    f->has_tree = true;
    // the clouds' own contexts are no longer needed: the forest holds every array the searches read
    for (auto &c : b->clouds) oct_destroy(c.ctx), c.ctx = nullptr;
    b->built = true;
    return OCT_OK;
}

int oct_batch_stage_times(const oct_batch *b, double ms[8]) {
    if (!b) return OCT_ERR_ARG;
    return oct_stage_times(b->forest, ms);
}

int oct_batch_info(const oct_batch *b, int cloud, size_t *n, oct_tree_info *info) {
    if (!b || cloud < 0 || (size_t)cloud >= b->clouds.size()) return OCT_ERR_ARG;
    if (n) *n = b->clouds[(size_t)cloud].n;
    if (info) *info = b->clouds[(size_t)cloud].info;
    return OCT_OK;
}

// queries grouped per cloud, each group in SFC (sorted index) order and padded to tiles of T slots
static void tile_queries(const oct_batch *b, const uint32_t *cloud, const uint32_t *center_idx, const std::vector<size_t> &ids,
                         uint32_t T, std::vector<uint32_t> &order, std::vector<uint32_t> &tiles) {
    const size_t C = b->clouds.size();
    std::vector<std::vector<uint32_t>> per(C);
    for (size_t q : ids) per[cloud[q]].push_back((uint32_t)q);
    order.clear(), tiles.clear();
    for (size_t c = 0; c < C; c++) {
        auto &v = per[c];
        if (v.empty()) continue;
        std::stable_sort(v.begin(), v.end(), [&](uint32_t x, uint32_t y) { return center_idx[x] < center_idx[y]; });
        for (size_t t = 0; t < v.size(); t += T) {
            const size_t cnt = std::min<size_t>(T, v.size() - t);
            tiles.push_back((uint32_t)c | (uint32_t)cnt << 20);
            for (size_t k = 0; k < T; k++) order.push_back(k < cnt ? v[t + k] : 0u);
        }
    }
}

static int check_queries(oct_batch *b, const uint32_t *cloud, const uint32_t *center_idx, size_t nq) {
    if (!b->built) return bfail(b, OCT_ERR_STATE, "batch search before oct_batch_build");
    if (nq && (!cloud || !center_idx)) return bfail(b, OCT_ERR_ARG, "null query arrays");
    if (nq >= 0xffffffffull) return bfail(b, OCT_ERR_ARG, "too many queries");
    for (size_t i = 0; i < nq; i++) {
        if (cloud[i] >= b->clouds.size()) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
        if (center_idx[i] >= b->clouds[cloud[i]].n) return bfail(b, OCT_ERR_ARG, "center index out of range");
    }
    return OCT_OK;
}

static int batch_stats(oct_batch *b, oct_search_stats *stats, Stage st) {
    unsigned long long h[6];
    oct_ctx *f = b->forest;
    BCHK(hipMemcpyAsync(h, f->stats.p, sizeof(h), hipMemcpyDeviceToHost, f->stream), "copy stats");
    BCHK(hipStreamSynchronize(f->stream), "stats");
    stats->nodes_classified = h[0], stats->points_tested = h[1], stats->rechecked = h[2], stats->total_neighbors = h[3];
    stats->class_slots = h[4], stats->test_slots = h[5];
    float ms = 0.f;
    hipEventElapsedTime(&ms, f->ev0[st], f->ev1[st]);
    stats->kernel_ms = ms;
    return OCT_OK;
}

int oct_batch_radius(oct_batch *b, int kernel, int algo, int precision, const uint32_t *cloud, const uint32_t *center_idx,
                     const double *radii, size_t nq, uint32_t *counts_out, oct_search_stats *stats) {
    if (!b) return OCT_ERR_ARG;
    if (kernel < 0 || kernel > 3 || algo < 1 || algo > 3 || (precision != 0 && precision != 1))
        return bfail(b, OCT_ERR_ARG, "unknown kernel, algorithm or precision");
    if (nq && (!radii || !counts_out)) return bfail(b, OCT_ERR_ARG, "null radii or counts");
    int r = check_queries(b, cloud, center_idx, nq);
    if (r) return r;
    if (nq == 0) return OCT_OK;
    BCHK(hipSetDevice(b->device), "set device");
    oct_ctx *f = b->forest;
    std::vector<size_t> ids(nq);
    std::iota(ids.begin(), ids.end(), (size_t)0);
    std::vector<uint32_t> order, tiles;
    tile_queries(b, cloud, center_idx, ids, 64, order, tiles);
    std::vector<uint32_t> gidx(nq);
    for (size_t i = 0; i < nq; i++) gidx[i] = (uint32_t)(b->clouds[cloud[i]].pt_off + center_idx[i]);
    BCHK(b->order.reserve(order.size()), "alloc");
    BCHK(b->tile_info.reserve(tiles.size()), "alloc");
    BCHK(b->qidx.reserve(nq), "alloc");
    BCHK(b->qrad.reserve(nq), "alloc");
    BCHK(b->counts.reserve(nq), "alloc");
    BCHK(hipMemcpyAsync(b->order.p, order.data(), order.size() * 4, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->tile_info.p, tiles.data(), tiles.size() * 4, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->qidx.p, gidx.data(), nq * 4, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->qrad.p, radii, nq * 8, hipMemcpyHostToDevice, f->stream), "copy");
    RadiusLaunch L;
    L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 0, L.group = 64, L.mx = false, L.remap = true;
    L.order = b->order.p, L.qidx = b->qidx.p, L.qrad = b->qrad.p, L.nq = order.size(), L.counts = b->counts.p;
    L.tile_info = b->tile_info.p, L.roots = b->roots.p, L.radtab = b->radtab.p;
    if (stats) {
        BCHK(f->stats.reserve(8), "alloc stats");
        BCHK(hipMemsetAsync(f->stats.p, 0, 8 * sizeof(unsigned long long), f->stream), "clear stats");
        L.stats = f->stats.p;
    }
    BCHK(launch_radius(f, L), "batch radius");
    BCHK(hipMemcpyAsync(counts_out, b->counts.p, nq * 4, hipMemcpyDeviceToHost, f->stream), "copy counts");
    BCHK(hipStreamSynchronize(f->stream), "batch radius");
    if (stats) return batch_stats(b, stats, ST_SEARCH);
    return OCT_OK;
}

int oct_batch_knn(oct_batch *b, int precision, const uint32_t *cloud, const uint32_t *center_idx, const uint32_t *k, size_t nq,
                  uint32_t kmax, int32_t *idx_out, double *d2_out, oct_search_stats *stats) {
    if (!b) return OCT_ERR_ARG;
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return bfail(b, OCT_ERR_ARG, "unknown precision");
    if (kmax == 0 || kmax > 1024) return bfail(b, OCT_ERR_ARG, "kmax must be 1 .. 1024");
    if (nq && (!k || !idx_out)) return bfail(b, OCT_ERR_ARG, "null k or outputs");
    int r = check_queries(b, cloud, center_idx, nq);
    if (r) return r;
    for (size_t i = 0; i < nq; i++)
        if (k[i] == 0 || k[i] > kmax) return bfail(b, OCT_ERR_ARG, "per-query k must be 1 .. kmax");
    if (nq == 0) return OCT_OK;
    BCHK(hipSetDevice(b->device), "set device");
    oct_ctx *f = b->forest;
    std::vector<uint32_t> gidx(nq);
    for (size_t i = 0; i < nq; i++) gidx[i] = (uint32_t)(b->clouds[cloud[i]].pt_off + center_idx[i]);
    BCHK(b->qidx.reserve(nq), "alloc");
    BCHK(b->kq.reserve(nq), "alloc");
    BCHK(b->kidx.reserve(nq * kmax), "alloc");
    if (d2_out) BCHK(b->kd2.reserve(nq * kmax), "alloc");
    BCHK(hipMemcpyAsync(b->qidx.p, gidx.data(), nq * 4, hipMemcpyHostToDevice, f->stream), "copy");
    BCHK(hipMemcpyAsync(b->kq.p, k, nq * 4, hipMemcpyHostToDevice, f->stream), "copy");
    if (stats) {
        BCHK(f->stats.reserve(8), "alloc stats");
        BCHK(hipMemsetAsync(f->stats.p, 0, 8 * sizeof(unsigned long long), f->stream), "clear stats");
    }
    // one launch per k size class (16, 32, ... 1024), every cloud in it
    std::vector<std::vector<size_t>> cls(7);
    for (size_t i = 0; i < nq; i++) {
        int c = 0;
        while (c < 6 && (16u << c) < k[i]) c++;
        cls[c].push_back(i);
    }
    float total_ms = 0.f;
    std::vector<std::vector<uint32_t>> orders(7), tiles(7);
    for (int c = 0; c < 7; c++) {
        if (cls[c].empty()) continue;
        const uint32_t kc = 16u << c;
        int opts[3];
        knn_lane_options(kc, opts);
        const uint32_t G = 64u / (uint32_t)opts[0];
        tile_queries(b, cloud, center_idx, cls[c], G, orders[c], tiles[c]);
        BCHK(b->order.reserve(orders[c].size()), "alloc");
        BCHK(b->tile_info.reserve(tiles[c].size()), "alloc");
        BCHK(hipMemcpyAsync(b->order.p, orders[c].data(), orders[c].size() * 4, hipMemcpyHostToDevice, f->stream), "copy");
        BCHK(hipMemcpyAsync(b->tile_info.p, tiles[c].data(), tiles[c].size() * 4, hipMemcpyHostToDevice, f->stream), "copy");
        KnnLaunch L;
        L.precision = precision;
        L.order = b->order.p, L.qidx = b->qidx.p, L.nq = orders[c].size(), L.k = kmax, L.kq = b->kq.p, L.kclass = kc;
        L.idx = b->kidx.p, L.d2 = d2_out ? b->kd2.p : nullptr;
        L.tile_info = b->tile_info.p, L.roots = b->roots.p, L.radtab = b->radtab.p, L.magtab = b->magtab.p;
        L.lanes = opts[0];
        if (stats) L.stats = f->stats.p;
        std::string err;
        hipError_t e = launch_knn(f, L, err);
        if (e != hipSuccess) return err.empty() ? bfail(b, OCT_ERR_HIP, "batch knn") : bfail(b, OCT_ERR_ARG, err);
        BCHK(hipStreamSynchronize(f->stream), "batch knn");  // order / tile staging is reused by the next class
        float ms = 0.f;
        hipEventElapsedTime(&ms, f->ev0[ST_KNN], f->ev1[ST_KNN]);
        total_ms += ms;
    }
    BCHK(hipMemcpyAsync(idx_out, b->kidx.p, nq * kmax * 4, hipMemcpyDeviceToHost, f->stream), "copy idx");
    if (d2_out) BCHK(hipMemcpyAsync(d2_out, b->kd2.p, nq * kmax * 8, hipMemcpyDeviceToHost, f->stream), "copy d2");
    BCHK(hipStreamSynchronize(f->stream), "batch knn");
    // forest indices -> the clouds' own sorted indices
    for (size_t i = 0; i < nq; i++) {
        const int32_t off = (int32_t)b->clouds[cloud[i]].pt_off;
        int32_t *row = idx_out + i * kmax;
        for (uint32_t t = 0; t < kmax; t++)
            if (row[t] >= 0) row[t] -= off;
    }
    if (stats) {
        int rr = batch_stats(b, stats, ST_KNN);
        stats->kernel_ms = total_ms;
        return rr;
    }
    return OCT_OK;
}

}  // extern "C"
