// Batches of independent clouds (no counterpart in the reference, whose LinearOctree holds one cloud;
// the reference harness runs its clouds one after another).
//
// oct_batch_add appends each cloud's points to one device buffer. Two builds lay the clouds out as
// one forest that the count and kNN kernels walk from any cloud's root:
//  - oct_batch_build: each cloud is encoded, sorted and built by a context of its own, exactly as a
//    single cloud, and its arrays are copied to point / node offsets (node offsets multiples of 8,
//    point ranges and child indices rebased), cloud-major;
//  - oct_batch_build_device: one device-driven sequence over all clouds at once (per-cloud boxes and
//    keys, a sort segmented by cloud, a level-synchronous forest build) with no host
//    synchronisation between its launches, optionally captured in a HIP graph on the first call and
//    replayed on later calls; nodes are level-major over the whole forest (pipeline.hip).
// Both give every cloud the same keys, permutation and tree as the single-cloud path, and both keep
// the child -> NodeExt rule ((first child - 1) / 8). A search over many clouds is then one launch per
// search kind: queries are grouped per cloud into wave tiles (64 slots for counts, 64 / lanes for
// kNN), and each tile names its cloud, which selects the root and that cloud's level radii.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "device.hpp"
#include "pipeline.hpp"

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
    oct::DevBuf<double> xyz_all;  // every cloud's points, packed x, y, z, in the order added
    size_t n_all = 0;
    bool built = false;
    int mode = 0;  // 1: per-cloud contexts (oct_batch_build), 2: device-driven (oct_batch_build_device)
    std::vector<double> h_rad, h_mag;
    std::vector<uint32_t> h_roots;
    oct::DevBuf<double> radtab, magtab;
    oct::DevBuf<uint32_t> roots, tile_info, order, qidx, kq, counts;
    oct::DevBuf<double> qrad;
    oct::DevBuf<int32_t> kidx;
    oct::DevBuf<double> kd2;
    // device-driven build
    bool dev_ready = false;
    uint32_t n_chunks = 0, n_tiles = 0, R = 0, cap = 0;
    oct::DevBuf<oct::Chunk> chunks;
    oct::DevBuf<oct::SortTile> stiles;
    oct::DevBuf<uint32_t> cloud_chunk0, cloud_off, cloud_enc, cloud_leaf;
    oct::DevBuf<oct::CloudGeom> geom;
    oct::DevBuf<double> partial;
    oct::DevBuf<uint32_t> hist, bases, ctrl, part, cl_internal, cl_depth, ncloud;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t lv[oct::kLvCount] = {0};
    double build_ms = 0.0;
    // search pass of the device pipeline: device-resident queries, grouped on the device each run
    struct Kind {
        uint32_t n = 0, nclass = 1;
        oct::DevBuf<uint32_t> cloud, idx, kq, gidx, vals0, vals1, hist, bases, lookback, ctrl, gstart, toff;
        oct::DevBuf<uint64_t> keys0, keys1;
        oct::DevBuf<uint32_t> order[oct::kKnnClasses], tile_info[oct::kKnnClasses];
        bool present[oct::kKnnClasses] = {};
        oct::QueryGroups Q;
    } rpass, kpass;
    oct::DevBuf<double> p_rad;
    oct::DevBuf<uint32_t> p_counts;
    oct::DevBuf<int32_t> p_kidx;
    oct::DevBuf<double> p_kd2;
    uint32_t p_kmax = 0;
    bool queries_set = false, results_ready = false;
    int knn_lanes[oct::kKnnClasses] = {0};  // lanes per kNN query per size class (0: the class's default)
    hipGraph_t pgraph = nullptr;
    hipGraphExec_t pexec = nullptr;
    int pkey[3 + oct::kKnnClasses] = {0};  // kernel, algo, precision, lanes: the captured pass's shape
    // host copy of the forest for per-cloud tree dumps (filled on first use after a build)
    bool dump_ready = false;
    std::vector<oct::NodeRec> h_nodes;
    std::vector<uint64_t> h_nkey;
    std::vector<uint32_t> h_child, h_ncloud;
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

// grow a device buffer to hold `need` elements, keeping its first `valid` (stream-ordered copy)
template <class T>
static hipError_t grow_keep(DevBuf<T> &b, size_t need, size_t valid, hipStream_t s) {
    if (need <= b.cap) return hipSuccess;
    size_t cap = b.cap ? b.cap : 1024;
    while (cap < need) cap *= 2;
    T *p = nullptr;
    hipError_t e = hipMalloc((void **)&p, cap * sizeof(T));
    if (e != hipSuccess) return e;
    if (b.p && valid) {
        if ((e = hipMemcpyAsync(p, b.p, valid * sizeof(T), hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    }
    if (b.p) (void)hipFree(b.p);
    b.p = p;
    b.cap = cap;
    return hipSuccess;
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
        if (e_ != hipSuccess) return bfail(b, e_ == hipErrorOutOfMemory ? OCT_ERR_NOMEM : OCT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BAPI(call, what)                                                                                   \
    do {                                                                                                   \
        int r_ = (call);                                                                                   \
        if (r_ != OCT_OK) return bfail(b, r_, std::string(what) + ": " + oct_last_error(ctx_for_error)); \
    } while (0)

// Synchronises the forest's stream and fails on a set device error word (a search kernel's full
// traversal stack).
static int forest_check(oct_batch *b) {
    const int r = oct_synchronize(b->forest);
    return r ? bfail(b, r, std::string("search: ") + oct_last_error(b->forest)) : OCT_OK;
}

static void drop_graph(oct_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->pexec) (void)hipGraphExecDestroy(b->pexec);
    if (b->pgraph) (void)hipGraphDestroy(b->pgraph);
    b->gexec = nullptr, b->graph = nullptr, b->pexec = nullptr, b->pgraph = nullptr;
}

static void release_kind(oct_batch::Kind &k) {
    k.cloud.release(), k.idx.release(), k.kq.release(), k.gidx.release(), k.vals0.release(), k.vals1.release();
    k.hist.release(), k.bases.release(), k.lookback.release(), k.ctrl.release(), k.gstart.release(), k.toff.release();
    k.keys0.release(), k.keys1.release();
    for (int c = 0; c < kKnnClasses; c++) k.order[c].release(), k.tile_info[c].release();
}

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
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->forest->stream);
    drop_graph(b);
    b->xyz_all.release();
    b->radtab.release(), b->magtab.release(), b->roots.release(), b->tile_info.release(), b->order.release();
    b->qidx.release(), b->kq.release(), b->counts.release(), b->qrad.release(), b->kidx.release(), b->kd2.release();
    b->chunks.release(), b->stiles.release(), b->cloud_chunk0.release(), b->cloud_off.release(), b->cloud_enc.release();
    b->cloud_leaf.release(), b->geom.release(), b->partial.release(), b->hist.release(), b->bases.release();
    b->ctrl.release(), b->part.release(), b->cl_internal.release(), b->cl_depth.release(), b->ncloud.release();
    release_kind(b->rpass), release_kind(b->kpass);
    b->p_rad.release(), b->p_counts.release(), b->p_kidx.release(), b->p_kd2.release();
    oct_destroy(b->forest);
    delete b;
    return OCT_OK;
}

const char *oct_batch_last_error(const oct_batch *b) { return b ? b->err.c_str() : "null batch"; }

// packed copy of n points (stride bytes apart) into the batch buffer at point offset `at`
static int put_points(oct_batch *b, const double *xyz, size_t n, size_t stride_bytes, size_t at) {
    if (!n) return OCT_OK;
    hipStream_t s = b->forest->stream;
    double *dst = b->xyz_all.p + 3 * at;
    if (stride_bytes == 24) {
        BCHK(hipMemcpyAsync(dst, xyz, n * 24, hipMemcpyHostToDevice, s), "copy points");
    } else {
        std::vector<double> packed(3 * n);
        const size_t sd = stride_bytes / 8;
        for (size_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) packed[3 * i + a] = xyz[sd * i + a];
        BCHK(hipMemcpyAsync(dst, packed.data(), n * 24, hipMemcpyHostToDevice, s), "copy points");
        BCHK(hipStreamSynchronize(s), "copy points");  // the staging vector ends here
    }
    return OCT_OK;
}

int oct_batch_add(oct_batch *b, const double *xyz, size_t n, size_t stride_bytes, int encoder, uint32_t max_points_leaf,
                  int *cloud_id) {
    if (!b) return OCT_ERR_ARG;
    // the device build sizes its tables and buffers for the clouds present at its first call
    if (b->built || b->mode || b->dev_ready) return bfail(b, OCT_ERR_STATE, "clouds cannot be added to a built batch");
    // the query grouping buffers are sized for the clouds present at oct_batch_queries
    if (b->queries_set) return bfail(b, OCT_ERR_STATE, "clouds cannot be added after oct_batch_queries");
    if (encoder != OCT_ENC_MORTON3D && encoder != OCT_ENC_HILBERT3D) return bfail(b, OCT_ERR_UNSUPPORTED, "encoder");
    if (max_points_leaf == 0) return bfail(b, OCT_ERR_ARG, "max_points_leaf must be positive");
    if (stride_bytes < 24 || stride_bytes % 8) return bfail(b, OCT_ERR_ARG, "stride_bytes must be >= 24 and a multiple of 8");
    if (n && !xyz) return bfail(b, OCT_ERR_ARG, "null point buffer");
    if (n >= (1ull << 30)) return bfail(b, OCT_ERR_ARG, "at most 2^30 - 1 points per cloud");
    if (b->n_all + n >= 0xffffffffull) return bfail(b, OCT_ERR_ARG, "batch beyond 2^32 - 1 points");
    if (b->clouds.size() >= (1u << 20) - 1) return bfail(b, OCT_ERR_ARG, "too many clouds in one batch");
    BCHK(hipSetDevice(b->device), "set device");
    BCHK(grow_keep(b->xyz_all, 3 * (b->n_all + n) + 3, 3 * b->n_all, b->forest->stream), "alloc points");
    int r = put_points(b, xyz, n, stride_bytes, b->n_all);
    if (r) return r;
    oct_batch::Cloud c;
    c.encoder = encoder, c.max_leaf = max_points_leaf, c.n = n, c.pt_off = b->n_all;
    b->n_all += n;
    b->clouds.push_back(c);
    if (cloud_id) *cloud_id = (int)b->clouds.size() - 1;
    return OCT_OK;
}

int oct_batch_set_points(oct_batch *b, int cloud, const double *xyz, size_t stride_bytes) {
    if (!b) return OCT_ERR_ARG;
    if (cloud < 0 || (size_t)cloud >= b->clouds.size()) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
    if (b->built && b->mode != 2) return bfail(b, OCT_ERR_STATE, "points of a batch built per cloud cannot change");
    if (stride_bytes < 24 || stride_bytes % 8) return bfail(b, OCT_ERR_ARG, "stride_bytes must be >= 24 and a multiple of 8");
    const auto &c = b->clouds[(size_t)cloud];
    if (c.n && !xyz) return bfail(b, OCT_ERR_ARG, "null point buffer");
    BCHK(hipSetDevice(b->device), "set device");
    int r = put_points(b, xyz, c.n, stride_bytes, c.pt_off);
    if (r) return r;
    BCHK(hipStreamSynchronize(b->forest->stream), "copy points");
    b->built = false;  // the forest no longer matches the points until the next oct_batch_build_device
    b->dump_ready = false;
    return OCT_OK;
}

int oct_batch_build(oct_batch *b) {
    if (!b) return OCT_ERR_ARG;
    if (b->built || b->mode) return bfail(b, OCT_ERR_STATE, "batch already built");
    BCHK(hipSetDevice(b->device), "set device");
    BCHK(hipStreamSynchronize(b->forest->stream), "points");
    // every cloud on its own: bbox, encode, sort, build (the single-cloud path, bit-identical)
    size_t pts = 0, nodes = 0;
    for (auto &c : b->clouds) {
        int r = oct_create(&c.ctx, b->device);
        if (r != OCT_OK) return bfail(b, r, "context of a cloud");
        oct_ctx *ctx_for_error = c.ctx;
        double cen[3], rad[3];
        BAPI(oct_set_points_dev(c.ctx, b->xyz_all.p + 3 * c.pt_off, c.n, 24), "points of a cloud");
        BAPI(oct_sort_points(c.ctx, c.encoder, cen, rad, nullptr, nullptr), "sort a cloud");
        BAPI(oct_build(c.ctx, c.max_leaf, &c.info), "build a cloud");
        BCHK(hipStreamSynchronize(c.ctx->stream), "build a cloud");
        c.node_off = nodes;
        pts += c.n;
        nodes += ((size_t)c.info.n_total + 7) / 8 * 8;
    }
    if (nodes >= 0xffffffffull) return bfail(b, OCT_ERR_ARG, "batch beyond 2^32 nodes");
    // the forest
    oct_ctx *f = b->forest;
    BCHK(f->p64.reserve(pts + 1), "alloc forest points");
    BCHK(f->p32.reserve(pts + 1), "alloc forest points");
    BCHK(f->keys[0].reserve(pts + 1), "alloc forest keys");
    BCHK(f->vals[0].reserve(pts + 1), "alloc forest keys");
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
            BCHK(hipMemcpyAsync(f->keys[0].p + c.pt_off, x->sorted_keys, c.n * 8, hipMemcpyDeviceToDevice, f->stream), "copy");
            BCHK(hipMemcpyAsync(f->vals[0].p + c.pt_off, x->perm, c.n * 4, hipMemcpyDeviceToDevice, f->stream), "copy");
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
        // the context's work must finish before the context (and its buffers) go
        BCHK(hipStreamSynchronize(f->stream), "assemble the forest");
        oct_destroy(c.ctx);
        b->clouds[i].ctx = nullptr;
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
    b->built = true;
    b->mode = 1;
    return OCT_OK;
}

// ------------------------------------------------------------------ device-driven build
// Work tables (chunks, sort tiles, cloud offsets) and every buffer the sequence touches; node
// storage at b->cap. Called before the first build and again when the node capacity grows.
static int device_setup(oct_batch *b) {
    oct_ctx *f = b->forest;
    hipStream_t s = f->stream;
    const uint32_t C = (uint32_t)b->clouds.size();
    const size_t n = b->n_all;
    if (!b->dev_ready) {
        std::vector<Chunk> ch;
        std::vector<SortTile> st;
        std::vector<uint32_t> c0(C + 1), off(C + 1), enc(C ? C : 1), leaf(C ? C : 1);
        size_t est = 0;
        for (uint32_t c = 0; c < C; c++) {
            const auto &cl = b->clouds[c];
            c0[c] = (uint32_t)ch.size();
            off[c] = (uint32_t)cl.pt_off;
            enc[c] = (uint32_t)cl.encoder, leaf[c] = cl.max_leaf;
            for (size_t p = 0; p < cl.n; p += kChunkPoints)
                ch.push_back(Chunk{(uint32_t)(cl.pt_off + p), (uint32_t)std::min<size_t>(kChunkPoints, cl.n - p), c, 0u});
            const uint32_t first = (uint32_t)st.size();
            for (size_t p = 0; p < cl.n; p += kSortTile)
                st.push_back(SortTile{(uint32_t)(cl.pt_off + p), (uint32_t)std::min<size_t>(kSortTile, cl.n - p), c, first});
            est += 64 + 8 * (cl.n / cl.max_leaf);
        }
        c0[C] = (uint32_t)ch.size();
        off[C] = (uint32_t)n;
        b->n_chunks = (uint32_t)ch.size();
        b->n_tiles = (uint32_t)st.size();
        b->R = C ? (C + 6) / 8 * 8 + 1 : 1;
        if (b->cap == 0) b->cap = (uint32_t)std::min<size_t>(b->R + est, 0xfffffff0ull);
        BCHK(b->chunks.reserve(ch.size() + 1), "alloc");
        BCHK(b->stiles.reserve(st.size() + 1), "alloc");
        BCHK(b->cloud_chunk0.reserve(C + 1), "alloc");
        BCHK(b->cloud_off.reserve(C + 1), "alloc");
        BCHK(b->cloud_enc.reserve(C + 1), "alloc");
        BCHK(b->cloud_leaf.reserve(C + 1), "alloc");
        if (!ch.empty()) BCHK(hipMemcpyAsync(b->chunks.p, ch.data(), ch.size() * sizeof(Chunk), hipMemcpyHostToDevice, s), "copy");
        if (!st.empty())
            BCHK(hipMemcpyAsync(b->stiles.p, st.data(), st.size() * sizeof(SortTile), hipMemcpyHostToDevice, s), "copy");
        BCHK(hipMemcpyAsync(b->cloud_chunk0.p, c0.data(), (C + 1) * 4, hipMemcpyHostToDevice, s), "copy");
        BCHK(hipMemcpyAsync(b->cloud_off.p, off.data(), (C + 1) * 4, hipMemcpyHostToDevice, s), "copy");
        BCHK(hipMemcpyAsync(b->cloud_enc.p, enc.data(), C * 4, hipMemcpyHostToDevice, s), "copy");
        BCHK(hipMemcpyAsync(b->cloud_leaf.p, leaf.data(), C * 4, hipMemcpyHostToDevice, s), "copy");
        BCHK(hipStreamSynchronize(s), "copy tables");  // the host vectors end here
        BCHK(b->geom.reserve(C + 1), "alloc");
        BCHK(b->radtab.reserve(66 * (size_t)(C + 1)), "alloc");
        BCHK(b->magtab.reserve(3 * (size_t)(C + 1)), "alloc");
        BCHK(b->roots.reserve(C + 1), "alloc");
        BCHK(b->partial.reserve(6 * (size_t)b->n_chunks + 6), "alloc");
        BCHK(b->hist.reserve((size_t)C * kSegBaseStride + 1), "alloc");
        BCHK(b->bases.reserve((size_t)C * kSegBaseStride + 1), "alloc");
        BCHK(f->lookback.reserve((size_t)kRadixPassCount * b->n_tiles * 256 + 1), "alloc");
        BCHK(b->ctrl.reserve(kCtrlWords), "alloc");
        BCHK(b->part.reserve(1024), "alloc");
        BCHK(b->cl_internal.reserve(C + 1), "alloc");
        BCHK(b->cl_depth.reserve(C + 1), "alloc");
        BCHK(f->keys[0].reserve(n + 1), "alloc keys");
        BCHK(f->keys[1].reserve(n + 1), "alloc keys");
        BCHK(f->vals[0].reserve(n + 1), "alloc keys");
        BCHK(f->vals[1].reserve(n + 1), "alloc keys");
        BCHK(f->p64.reserve(n + 1), "alloc points");
        BCHK(f->p32.reserve(n + 1), "alloc points");
        b->dev_ready = true;
    }
    const size_t cap = b->cap;
    BCHK(f->nodes.reserve(cap), "alloc nodes");
    BCHK(f->nkey.reserve(cap), "alloc nodes");
    BCHK(f->child.reserve(cap + 1), "alloc nodes");
    BCHK(f->flags.reserve(cap), "alloc nodes");
    BCHK(f->amax.reserve(cap), "alloc nodes");
    BCHK(b->ncloud.reserve(cap), "alloc nodes");
    BCHK(f->next.reserve(cap / 8 + 1), "alloc nodes");
    return OCT_OK;
}

static ForestBuild forest_args(oct_batch *b) {
    oct_ctx *f = b->forest;
    ForestBuild F;
    F.n = b->n_all, F.C = (uint32_t)b->clouds.size(), F.R = b->R, F.cap = b->cap;
    F.xyz = b->xyz_all.p;
    F.chunks = b->chunks.p, F.nchunks = b->n_chunks;
    F.cloud_chunk0 = b->cloud_chunk0.p, F.cloud_off = b->cloud_off.p, F.cloud_enc = b->cloud_enc.p, F.cloud_leaf = b->cloud_leaf.p;
    F.tiles = b->stiles.p, F.ntiles = b->n_tiles;
    F.geom = b->geom.p, F.radtab = b->radtab.p, F.magtab = b->magtab.p, F.roots = b->roots.p;
    F.partial = b->partial.p;
    F.keys[0] = f->keys[0].p, F.keys[1] = f->keys[1].p, F.vals[0] = f->vals[0].p, F.vals[1] = f->vals[1].p;
    F.hist = b->hist.p, F.bases = b->bases.p, F.lookback = f->lookback.p, F.ctrl = b->ctrl.p, F.part = b->part.p;
    F.cl_internal = b->cl_internal.p, F.cl_depth = b->cl_depth.p;
    F.p64 = f->p64.p, F.p32 = f->p32.p;
    F.nodes = f->nodes.p, F.nkey = f->nkey.p, F.child = f->child.p, F.flags = f->flags.p, F.ncloud = b->ncloud.p;
    F.amax = f->amax.p, F.ext = f->next.p;
    return F;
}

// lanes per query of kNN class c (the autotuned or set value, else the class's default)
static int class_lanes(const oct_batch *b, int c) {
    int opts[3];
    knn_lane_options(16u << c, opts);
    return b->knn_lanes[c] > 0 ? b->knn_lanes[c] : opts[0];
}

// grouping shape of the kNN queries for the current lanes per class
static void knn_shape(oct_batch *b) {
    const uint32_t C = (uint32_t)b->clouds.size();
    for (int c = 0; c < kKnnClasses; c++) {
        b->kpass.Q.G[c] = 64u / (uint32_t)class_lanes(b, c);
        b->kpass.Q.slots[c] = group_slots(b->kpass.n, C, b->kpass.Q.G[c]);
    }
}

// the search pass: radius queries grouped and counted (work-queue count kernel), kNN queries grouped
// and searched per size class (work-queue kNN kernel, which writes indices local to each query's
// cloud). only_class >= 0: just that kNN class (autotuning), timed with ev0 / ev1 when given.
static hipError_t enqueue_search(oct_batch *b, int kernel, int algo, int precision, std::string &err, int only_class = -1,
                                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    oct_ctx *f = b->forest;
    hipStream_t s = f->stream;
    const uint32_t C = (uint32_t)b->clouds.size();
    hipError_t e;
    if (b->rpass.n && only_class < 0) {
        if ((e = enqueue_query_groups(s, b->rpass.Q, C, b->cloud_off.p)) != hipSuccess) return e;
        RadiusLaunch L;
        L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 0, L.group = 64, L.remap = true;
        L.order = b->rpass.Q.order[0], L.qidx = b->rpass.Q.gidx, L.qrad = b->p_rad.p, L.nq = b->rpass.Q.slots[0];
        L.counts = b->p_counts.p, L.tile_info = b->rpass.Q.tile_info[0], L.roots = b->roots.p, L.radtab = b->radtab.p;
        L.queue_tiles = b->rpass.Q.ctrl + kQCtrlTiles, L.queue_next = b->rpass.Q.ctrl + kQCtrlQueue;
        if ((e = launch_radius(f, L)) != hipSuccess) return e;
    }
    if (b->kpass.n) {
        if ((e = enqueue_query_groups(s, b->kpass.Q, C, b->cloud_off.p)) != hipSuccess) return e;
        for (int c = 0; c < kKnnClasses; c++) {
            if (!b->kpass.present[c] || (only_class >= 0 && c != only_class)) continue;
            KnnLaunch L;
            L.precision = precision;
            L.order = b->kpass.Q.order[c], L.qidx = b->kpass.Q.gidx, L.nq = b->kpass.Q.slots[c], L.k = b->p_kmax, L.kq = b->kpass.Q.kq;
            L.kclass = 16u << c, L.lanes = class_lanes(b, c);
            L.idx = b->p_kidx.p, L.d2 = b->p_kd2.p;
            L.tile_info = b->kpass.Q.tile_info[c], L.roots = b->roots.p, L.radtab = b->radtab.p, L.magtab = b->magtab.p;
            L.queue_tiles = b->kpass.Q.ctrl + kQCtrlTiles + c, L.queue_next = b->kpass.Q.ctrl + kQCtrlQueue + c;
            if (only_class < 0) L.cloud_off = b->cloud_off.p;  // indices local to each query's cloud
            if (ev0) (void)hipEventRecord(ev0, s);
            if ((e = launch_knn(f, L, err)) != hipSuccess) return e;
            if (ev1) (void)hipEventRecord(ev1, s);
        }
    }
    return hipGetLastError();
}

// Device build (and with `search`, the search pass) of every cloud: captured in a graph on first use
// when use_graph, else launched directly; grows the node storage and runs again on overflow; fills
// the per-cloud tree info.
static int run_device(oct_batch *b, int use_graph, bool search, int kernel, int algo, int precision, double *ms_out) {
    oct_ctx *f = b->forest;
    hipStream_t s = f->stream;
    BCHK(hipStreamSynchronize(s), "points");
    const uint32_t C = (uint32_t)b->clouds.size();
    std::vector<uint32_t> ctrl(kCtrlWords), rctrl(kCtrlWords), kctrl(kCtrlWords), cint(C), cdep(C);
    int key[3 + kKnnClasses] = {kernel, algo, precision};
    for (int c = 0; c < kKnnClasses; c++) key[3 + c] = class_lanes(b, c);
    if (search && b->pexec && !std::equal(key, key + 3 + kKnnClasses, b->pkey)) {
        (void)hipGraphExecDestroy(b->pexec), (void)hipGraphDestroy(b->pgraph);
        b->pexec = nullptr, b->pgraph = nullptr;
    }
    f->stage_events_off = true;
    struct Restore {
        oct_ctx *f;
        ~Restore() { f->stage_events_off = false; }
    } restore{f};
    for (int attempt = 0;; attempt++) {
        int r = device_setup(b);
        if (r) return r;
        hipGraph_t *gp = search ? &b->pgraph : &b->graph;
        hipGraphExec_t *xp = search ? &b->pexec : &b->gexec;
        std::string err;
        if (use_graph && !*xp) {
            BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
            hipError_t e = enqueue_forest_build(s, forest_args(b));
            if (e == hipSuccess && search) e = enqueue_search(b, kernel, algo, precision, err);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(s, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (!err.empty()) return bfail(b, OCT_ERR_ARG, err);
                return bfail(b, OCT_ERR_HIP, std::string("capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
            }
            *gp = g;
            BCHK(hipGraphInstantiate(xp, *gp, nullptr, nullptr, 0), "instantiate the graph");
            if (search) std::copy(key, key + 3 + kKnnClasses, b->pkey);
        }
        BCHK(hipEventRecord(f->ev0[ST_BUILD], s), "event");
        if (use_graph) {
            BCHK(hipGraphLaunch(*xp, s), "launch the graph");
        } else {
            hipError_t e = enqueue_forest_build(s, forest_args(b));
            if (e == hipSuccess && search) e = enqueue_search(b, kernel, algo, precision, err);
            if (!err.empty()) return bfail(b, OCT_ERR_ARG, err);
            BCHK(e, "device build");
        }
        BCHK(hipEventRecord(f->ev1[ST_BUILD], s), "event");
        BCHK(hipMemcpyAsync(ctrl.data(), b->ctrl.p, kCtrlWords * 4, hipMemcpyDeviceToHost, s), "copy status");
        if (search && b->rpass.n) BCHK(hipMemcpyAsync(rctrl.data(), b->rpass.ctrl.p, kCtrlWords * 4, hipMemcpyDeviceToHost, s), "copy status");
        if (search && b->kpass.n) BCHK(hipMemcpyAsync(kctrl.data(), b->kpass.ctrl.p, kCtrlWords * 4, hipMemcpyDeviceToHost, s), "copy status");
        BCHK(hipMemcpyAsync(cint.data(), b->cl_internal.p, C * 4, hipMemcpyDeviceToHost, s), "copy status");
        BCHK(hipMemcpyAsync(cdep.data(), b->cl_depth.p, C * 4, hipMemcpyDeviceToHost, s), "copy status");
        BCHK(hipStreamSynchronize(s), "device build");
        if (search)
            if (int fr = forest_check(b)) return fr;
        if (ctrl[kCtrlSortErr] || rctrl[kCtrlSortErr] || kctrl[kCtrlSortErr])
            return bfail(b, OCT_ERR_STATE, "sort: look-back wait exceeded its bound");
        if (ctrl[kCtrlBuildErr] & kBuildErrDepth) return bfail(b, OCT_ERR_STATE, "build: split below the maximum depth");
        if (ctrl[kCtrlBuildErr] & kBuildErrCapacity) {
            if (attempt >= 8 || b->cap >= 0x7fffffffu) return bfail(b, OCT_ERR_NOMEM, "build: node capacity");
            b->cap = (uint32_t)std::min<uint64_t>(2ull * b->cap, 0xfffffff0ull);
            drop_graph(b);  // node buffers move
            continue;
        }
        break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, f->ev0[ST_BUILD], f->ev1[ST_BUILD]);
    b->build_ms = ms;
    if (ms_out) *ms_out = ms;
    for (uint32_t l = 0; l < kLvCount; l++) b->lv[l] = ctrl[kCtrlLv + l];
    b->h_rad.assign(66 * (size_t)C, 0.0);
    b->h_mag.assign(3 * (size_t)C, 0.0);
    BCHK(hipMemcpyAsync(b->h_rad.data(), b->radtab.p, 66 * (size_t)C * 8, hipMemcpyDeviceToHost, s), "copy radii");
    BCHK(hipMemcpyAsync(b->h_mag.data(), b->magtab.p, 3 * (size_t)C * 8, hipMemcpyDeviceToHost, s), "copy radii");
    BCHK(hipStreamSynchronize(s), "copy radii");
    for (uint32_t c = 0; c < C; c++) {
        auto &cl = b->clouds[c];
        oct_tree_info &t = cl.info;
        t.n_internal = cint[c];
        t.n_total = 1 + 8 * cint[c];
        t.n_leaf = t.n_total - t.n_internal;
        t.max_depth_seen = cdep[c];
        t.n_levels = cdep[c] + 1;
        t.max_points_leaf = cl.max_leaf;
        t.build_ms = ms;  // the whole batch's device time
    }
    f->n = b->n_all;
    f->has_tree = true;
    b->built = true;
    b->mode = 2;
    b->dump_ready = false;
    b->results_ready = search;
    return OCT_OK;
}

int oct_batch_build_device(oct_batch *b, int use_graph, double *ms_out) {
    if (!b) return OCT_ERR_ARG;
    if (b->mode == 1) return bfail(b, OCT_ERR_STATE, "batch already built per cloud");
    if (b->clouds.empty()) return bfail(b, OCT_ERR_STATE, "empty batch");
    BCHK(hipSetDevice(b->device), "set device");
    return run_device(b, use_graph, false, 0, 0, 0, ms_out);
}

// device copies of one query kind and its grouping buffers
static int setup_kind(oct_batch *b, oct_batch::Kind &K, const uint32_t *cloud, const uint32_t *idx, const uint32_t *k, size_t n) {
    hipStream_t s = b->forest->stream;
    const uint32_t C = (uint32_t)b->clouds.size();
    K.n = (uint32_t)n;
    K.nclass = k ? kKnnClasses : 1;
    const size_t n1 = n + 1, tiles = (n + kSortTile - 1) / kSortTile;
    BCHK(K.cloud.reserve(n1), "alloc queries");
    BCHK(K.idx.reserve(n1), "alloc queries");
    BCHK(K.gidx.reserve(n1), "alloc queries");
    BCHK(K.vals0.reserve(n1), "alloc queries");
    BCHK(K.vals1.reserve(n1), "alloc queries");
    BCHK(K.keys0.reserve(n1), "alloc queries");
    BCHK(K.keys1.reserve(n1), "alloc queries");
    BCHK(K.hist.reserve(kRadixPassCount * 256), "alloc queries");
    BCHK(K.bases.reserve(kRadixPassCount * 256), "alloc queries");
    BCHK(K.lookback.reserve(5 * tiles * 256 + 1), "alloc queries");
    BCHK(K.ctrl.reserve(kCtrlWords), "alloc queries");
    BCHK(K.gstart.reserve((size_t)K.nclass * C + 1), "alloc queries");
    BCHK(K.toff.reserve((size_t)K.nclass * C + 1), "alloc queries");
    if (k) BCHK(K.kq.reserve(n1), "alloc queries");
    for (uint32_t c = 0; c < K.nclass; c++) {
        BCHK(K.order[c].reserve((size_t)group_slots((uint32_t)n, C, 64) + 64), "alloc queries");
        BCHK(K.tile_info[c].reserve(n + C + 1), "alloc queries");
    }
    if (n) {
        BCHK(hipMemcpyAsync(K.cloud.p, cloud, n * 4, hipMemcpyHostToDevice, s), "copy queries");
        BCHK(hipMemcpyAsync(K.idx.p, idx, n * 4, hipMemcpyHostToDevice, s), "copy queries");
        if (k) BCHK(hipMemcpyAsync(K.kq.p, k, n * 4, hipMemcpyHostToDevice, s), "copy queries");
    }
    for (int c = 0; c < kKnnClasses; c++) K.present[c] = false;
    if (k)
        for (size_t i = 0; i < n; i++) K.present[knn_class_of(k[i])] = true;
    QueryGroups &Q = K.Q;
    Q.n = K.n, Q.nclass = K.nclass;
    Q.cloud = K.cloud.p, Q.idx = K.idx.p, Q.kq = k ? K.kq.p : nullptr;
    Q.keys[0] = K.keys0.p, Q.keys[1] = K.keys1.p, Q.vals[0] = K.vals0.p, Q.vals[1] = K.vals1.p;
    Q.gidx = K.gidx.p, Q.hist = K.hist.p, Q.bases = K.bases.p, Q.lookback = K.lookback.p, Q.ctrl = K.ctrl.p;
    Q.gstart = K.gstart.p, Q.toff = K.toff.p;
    for (uint32_t c = 0; c < (uint32_t)kKnnClasses; c++) {
        Q.order[c] = c < K.nclass ? K.order[c].p : nullptr;
        Q.tile_info[c] = c < K.nclass ? K.tile_info[c].p : nullptr;
        Q.G[c] = 64;
        Q.slots[c] = c < K.nclass ? group_slots((uint32_t)n, C, 64) : 0u;
    }
    return OCT_OK;
}

int oct_batch_queries(oct_batch *b, const uint32_t *r_cloud, const uint32_t *r_idx, const double *r_radius, size_t nr,
                      const uint32_t *k_cloud, const uint32_t *k_idx, const uint32_t *k_k, size_t nk, uint32_t kmax) {
    if (!b) return OCT_ERR_ARG;
    if (b->mode == 1) return bfail(b, OCT_ERR_STATE, "the search pass runs on batches built on the device");
    if (b->clouds.empty()) return bfail(b, OCT_ERR_STATE, "empty batch");
    if (nr >= (1ull << 30) || nk >= (1ull << 30)) return bfail(b, OCT_ERR_ARG, "at most 2^30 - 1 queries of each kind");
    if (nr && (!r_cloud || !r_idx || !r_radius)) return bfail(b, OCT_ERR_ARG, "null radius query arrays");
    if (nk && (!k_cloud || !k_idx || !k_k)) return bfail(b, OCT_ERR_ARG, "null kNN query arrays");
    if (nk && (kmax == 0 || kmax > 1024)) return bfail(b, OCT_ERR_ARG, "kmax must be 1 .. 1024");
    if (nk && nk * (size_t)kmax >= (1ull << 40)) return bfail(b, OCT_ERR_ARG, "kNN output too large");
    const size_t C = b->clouds.size();
    for (size_t i = 0; i < nr; i++) {
        if (r_cloud[i] >= C) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
        if (r_idx[i] >= b->clouds[r_cloud[i]].n) return bfail(b, OCT_ERR_ARG, "center index out of range");
    }
    for (size_t i = 0; i < nk; i++) {
        if (k_cloud[i] >= C) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
        if (k_idx[i] >= b->clouds[k_cloud[i]].n) return bfail(b, OCT_ERR_ARG, "center index out of range");
        if (k_k[i] == 0 || k_k[i] > kmax) return bfail(b, OCT_ERR_ARG, "per-query k must be 1 .. kmax");
    }
    BCHK(hipSetDevice(b->device), "set device");
    hipStream_t s = b->forest->stream;
    BCHK(hipStreamSynchronize(s), "queries");
    drop_graph(b);  // query buffers move
    int r = setup_kind(b, b->rpass, r_cloud, r_idx, nullptr, nr);
    if (r) return r;
    r = setup_kind(b, b->kpass, k_cloud, k_idx, k_k, nk);
    if (r) return r;
    knn_shape(b);
    BCHK(b->p_rad.reserve(nr + 1), "alloc queries");
    BCHK(b->p_counts.reserve(nr + 1), "alloc results");
    BCHK(b->p_kidx.reserve(nk * (size_t)kmax + 1), "alloc results");
    BCHK(b->p_kd2.reserve(nk * (size_t)kmax + 1), "alloc results");
    if (nr) BCHK(hipMemcpyAsync(b->p_rad.p, r_radius, nr * 8, hipMemcpyHostToDevice, s), "copy queries");
    BCHK(hipStreamSynchronize(s), "copy queries");
    b->p_kmax = nk ? kmax : 0;
    b->queries_set = true;
    b->results_ready = false;
    return OCT_OK;
}

int oct_batch_pipeline(oct_batch *b, int kernel, int algo, int precision, int use_graph, double *ms) {
    if (!b) return OCT_ERR_ARG;
    if (!b->queries_set) return bfail(b, OCT_ERR_STATE, "oct_batch_pipeline before oct_batch_queries");
    if (b->mode == 1) return bfail(b, OCT_ERR_STATE, "batch already built per cloud");
    if (kernel < 0 || kernel > 3 || algo < 1 || algo > 3 || (precision != 0 && precision != 1))
        return bfail(b, OCT_ERR_ARG, "unknown kernel, algorithm or precision");
    BCHK(hipSetDevice(b->device), "set device");
    return run_device(b, use_graph, true, kernel, algo, precision, ms);
}

int oct_batch_pipeline_results(oct_batch *b, uint32_t *counts, int32_t *idx, double *d2) {
    if (!b) return OCT_ERR_ARG;
    if (!b->results_ready) return bfail(b, OCT_ERR_STATE, "no search pass results");
    BCHK(hipSetDevice(b->device), "set device");
    hipStream_t s = b->forest->stream;
    if (counts && b->rpass.n) BCHK(hipMemcpyAsync(counts, b->p_counts.p, (size_t)b->rpass.n * 4, hipMemcpyDeviceToHost, s), "copy counts");
    const size_t ke = (size_t)b->kpass.n * b->p_kmax;
    if (idx && ke) BCHK(hipMemcpyAsync(idx, b->p_kidx.p, ke * 4, hipMemcpyDeviceToHost, s), "copy kNN");
    if (d2 && ke) BCHK(hipMemcpyAsync(d2, b->p_kd2.p, ke * 8, hipMemcpyDeviceToHost, s), "copy kNN");
    BCHK(hipStreamSynchronize(s), "copy results");
    return OCT_OK;
}

// Per kNN size class present in the queries: every lanes-per-query shape offered for it is timed on
// the batch's own queries of that class (the forest as last built), the fastest kept.
int oct_batch_autotune_knn(oct_batch *b, int precision, int lanes_out[7], double ms_out[21]) {
    if (!b) return OCT_ERR_ARG;
    if (!b->queries_set || !b->built || b->mode != 2) return bfail(b, OCT_ERR_STATE, "autotune needs queries and a device-built batch");
    if (precision != 0 && precision != 1) return bfail(b, OCT_ERR_ARG, "unknown precision");
    BCHK(hipSetDevice(b->device), "set device");
    oct_ctx *f = b->forest;
    hipStream_t s = f->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BCHK(hipEventCreate(&e0), "event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return bfail(b, OCT_ERR_HIP, "event");
    }
    f->stage_events_off = true;
    int rc = OCT_OK;
    for (int c = 0; c < kKnnClasses && rc == OCT_OK; c++) {
        if (ms_out)
            for (int o = 0; o < 3; o++) ms_out[3 * c + o] = -1.0;
        if (!b->kpass.present[c]) continue;
        int opts[3];
        const int no = knn_lane_options(16u << c, opts);
        int best = opts[0];
        float best_ms = 0.f;
        for (int o = 0; o < no; o++) {
            b->knn_lanes[c] = opts[o];
            knn_shape(b);
            std::string err;
            hipError_t e = enqueue_search(b, 0, 1, precision, err, c, e0, e1);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                rc = bfail(b, err.empty() ? OCT_ERR_HIP : OCT_ERR_ARG, err.empty() ? std::string("autotune: ") + hipGetErrorString(e) : err);
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms_out) ms_out[3 * c + o] = ms;
            if (o == 0 || ms < best_ms) best = opts[o], best_ms = ms;
        }
        b->knn_lanes[c] = best;
    }
    f->stage_events_off = false;
    knn_shape(b);
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    if (lanes_out)
        for (int c = 0; c < kKnnClasses; c++) lanes_out[c] = class_lanes(b, c);
    b->results_ready = false;  // the timed runs wrote partial results
    return rc;
}

// sorted keys and permutation (indices into the cloud's points as added) of one cloud
int oct_batch_cloud_sorted(oct_batch *b, int cloud, uint64_t *keys_out, uint32_t *perm_out) {
    if (!b) return OCT_ERR_ARG;
    if (!b->built) return bfail(b, OCT_ERR_STATE, "batch not built");
    if (cloud < 0 || (size_t)cloud >= b->clouds.size()) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
    const auto &c = b->clouds[(size_t)cloud];
    if (!c.n) return OCT_OK;
    BCHK(hipSetDevice(b->device), "set device");
    oct_ctx *f = b->forest;
    if (keys_out) BCHK(hipMemcpyAsync(keys_out, f->keys[0].p + c.pt_off, c.n * 8, hipMemcpyDeviceToHost, f->stream), "copy keys");
    if (perm_out) BCHK(hipMemcpyAsync(perm_out, f->vals[0].p + c.pt_off, c.n * 4, hipMemcpyDeviceToHost, f->stream), "copy perm");
    BCHK(hipStreamSynchronize(f->stream), "copy");
    if (perm_out && b->mode == 2)  // the device build sorts global positions
        for (size_t i = 0; i < c.n; i++) perm_out[i] -= (uint32_t)c.pt_off;
    return OCT_OK;
}

// One cloud's tree in the reference's numbering, as oct_get_tree (device-built batches: the cloud's
// nodes taken in forest order are level-major, key-minor).
int oct_batch_cloud_tree(oct_batch *b, int cloud, uint64_t *prefixes, uint32_t *offsets, uint64_t *ranges, double *centers,
                         double *radii) {
    if (!b) return OCT_ERR_ARG;
    if (!b->built || b->mode != 2) return bfail(b, OCT_ERR_STATE, "tree dumps need oct_batch_build_device");
    if (cloud < 0 || (size_t)cloud >= b->clouds.size()) return bfail(b, OCT_ERR_ARG, "cloud id out of range");
    BCHK(hipSetDevice(b->device), "set device");
    oct_ctx *f = b->forest;
    const uint32_t nt = b->lv[kLvCount - 1];
    if (!b->dump_ready) {
        b->h_nodes.resize(nt), b->h_nkey.resize(nt), b->h_child.resize(nt), b->h_ncloud.resize(nt);
        BCHK(hipMemcpyAsync(b->h_nodes.data(), f->nodes.p, (size_t)nt * sizeof(NodeRec), hipMemcpyDeviceToHost, f->stream), "copy");
        BCHK(hipMemcpyAsync(b->h_nkey.data(), f->nkey.p, (size_t)nt * 8, hipMemcpyDeviceToHost, f->stream), "copy");
        BCHK(hipMemcpyAsync(b->h_child.data(), f->child.p, (size_t)nt * 4, hipMemcpyDeviceToHost, f->stream), "copy");
        BCHK(hipMemcpyAsync(b->h_ncloud.data(), b->ncloud.p, (size_t)nt * 4, hipMemcpyDeviceToHost, f->stream), "copy");
        BCHK(hipStreamSynchronize(f->stream), "copy");
        b->dump_ready = true;
    }
    const auto &c = b->clouds[(size_t)cloud];
    std::vector<uint32_t> ids;
    ids.reserve(c.info.n_total);
    for (uint32_t u = 0; u < nt; u++)
        if (b->h_ncloud[u] == (uint32_t)cloud) ids.push_back(u);
    if (ids.size() != c.info.n_total) return bfail(b, OCT_ERR_STATE, "forest node count of the cloud is inconsistent");
    std::vector<uint32_t> local(nt, 0u);
    for (size_t i = 0; i < ids.size(); i++) local[ids[i]] = (uint32_t)i;
    uint32_t level = 0;
    for (size_t i = 0; i < ids.size(); i++) {
        const uint32_t u = ids[i];
        while (level + 1 < kLvCount && b->lv[level + 1] <= u) level++;
        const NodeRec &r = b->h_nodes[u];
        if (prefixes) prefixes[i] = encode_placeholder(b->h_nkey[u], level);
        if (offsets) offsets[i] = b->h_child[u] ? local[b->h_child[u]] : 0u;
        if (ranges) ranges[2 * i] = r.first - c.pt_off, ranges[2 * i + 1] = r.last - c.pt_off;
        if (centers) centers[3 * i] = r.cx, centers[3 * i + 1] = r.cy, centers[3 * i + 2] = r.cz;
    }
    if (radii)
        for (int k = 0; k < 66; k++) radii[k] = b->h_rad[66 * (size_t)cloud + k];
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
    (void)hipEventElapsedTime(&ms, f->ev0[st], f->ev1[st]);
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
    L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 0, L.group = 64, L.remap = true;
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
    if (int fr = forest_check(b)) return fr;
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
        (void)hipEventElapsedTime(&ms, f->ev0[ST_KNN], f->ev1[ST_KNN]);
        total_ms += ms;
    }
    BCHK(hipMemcpyAsync(idx_out, b->kidx.p, nq * kmax * 4, hipMemcpyDeviceToHost, f->stream), "copy idx");
    if (d2_out) BCHK(hipMemcpyAsync(d2_out, b->kd2.p, nq * kmax * 8, hipMemcpyDeviceToHost, f->stream), "copy d2");
    BCHK(hipStreamSynchronize(f->stream), "batch knn");
    if (int fr = forest_check(b)) return fr;
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
