// C ABI (include/octree_amd.h): context management, argument checking, host<->device staging, and
// the sequencing of the gfx950 kernels. Every HIP call is checked; errors become negative status
// codes with a message in oct_last_error().
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <mutex>
#include <numeric>
#include <string>

#include "device.hpp"

using namespace oct;

namespace oct {
void stage_begin(oct_ctx *ctx, Stage s) {
    if (!ctx->stage_events_off) (void)hipEventRecord(ctx->ev0[s], ctx->stream);
}
void stage_end(oct_ctx *ctx, Stage s) {
    if (ctx->stage_events_off) return;
    (void)hipEventRecord(ctx->ev1[s], ctx->stream);
    ctx->ev_used[s] = true;
}
}  // namespace oct

namespace {

int fail(oct_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hfail(oct_ctx *ctx, hipError_t e, const char *where, const std::string &extra = "") {
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    if (!extra.empty()) m += " (" + extra + ")";
    return fail(ctx, e == hipErrorOutOfMemory ? OCT_ERR_NOMEM : OCT_ERR_HIP, m);
}

#define HCHK(expr, where)                                  \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hfail(ctx, _e, where); \
    } while (0)

bool valid_encoder(int e) { return e == OCT_ENC_MORTON3D || e == OCT_ENC_HILBERT3D; }

// Reads the device error word after the stream's work so far (synchronises); a set bit fails the
// call with OCT_ERR_STATE and the word is cleared for the next call.
int check_device_error(oct_ctx *ctx) {
    uint32_t w = 0;
    HCHK(hipMemcpyAsync(&w, ctx->derr.p, 4, hipMemcpyDeviceToHost, ctx->stream), "copy error word");
    HCHK(hipStreamSynchronize(ctx->stream), "copy error word");
    if (!w) return OCT_OK;
    HCHK(hipMemsetAsync(ctx->derr.p, 0, 4, ctx->stream), "clear error word");
    HCHK(hipStreamSynchronize(ctx->stream), "clear error word");
    if (w & kErrStack) return fail(ctx, OCT_ERR_STATE, "traversal stack full: a search dropped a pending node");
    return fail(ctx, OCT_ERR_STATE, "device error word set");
}

__global__ void k_pad(const double *__restrict__ xyz, double4 *__restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
}

__global__ void k_unstride(const double *__restrict__ src, size_t stride_d, double *__restrict__ xyz, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        xyz[3 * i] = src[stride_d * i];
        xyz[3 * i + 1] = src[stride_d * i + 1];
        xyz[3 * i + 2] = src[stride_d * i + 2];
    }
}

// Interleaves three coordinate arrays (x[0..n), y[0..n), z[0..n) back to back in src) into packed xyz.
__global__ void k_interleave(const double *__restrict__ src, double *__restrict__ xyz, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        xyz[3 * i] = src[i];
        xyz[3 * i + 1] = src[n + i];
        xyz[3 * i + 2] = src[2 * n + i];
    }
}

// Checksum of kNN result rows (nq rows of k int32 indices and FP32 squared distances): the wrapping
// sum over entries of a 64-bit mix of (position, index, distance bits) -- equal sums mean equal
// arrays up to a 2^-64 chance -- plus the rows whose distances are not ascending or whose indices
// lie outside [0, n), and the sum of the indices. One wave-reduced atomic per wave.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__global__ void k_knn_checksum(const int32_t *__restrict__ idx, const float *__restrict__ d2, size_t nq, uint32_t k,
                               uint32_t n, unsigned long long *__restrict__ out) {
    uint64_t h = 0, bad = 0, isum = 0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (size_t)gridDim.x * blockDim.x) {
        const int32_t *ri = idx + q * k;
        const float *rd = d2 + q * k;
        bool ok = true;
        float prev = -INFINITY;
        for (uint32_t j = 0; j < k; j++) {
            const int32_t i = ri[j];
            const float d = rd[j];
            ok &= (i >= 0 && (uint32_t)i < n && !(d < prev));
            prev = d;
            isum += (uint32_t)i;
            h += mix64((q * k + j) * 0x100000001b3ull ^ (((uint64_t)(uint32_t)i << 32) | __float_as_uint(d)));
        }
        bad += !ok;
    }
    for (int o = 32; o > 0; o >>= 1) {
        h += __shfl_xor(h, o);
        bad += __shfl_xor(bad, o);
        isum += __shfl_xor(isum, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], (unsigned long long)h);
        atomicAdd(&out[1], (unsigned long long)bad);
        atomicAdd(&out[2], (unsigned long long)isum);
    }
}

// Streaming copy, one 16-byte word per thread over the whole grid, non-temporal loads and stores: the
// fastest plain copy measured on this device (6.2-6.5 TB/s for 1 GiB against 4.3-4.8 TB/s for
// grid-stride loops of 4-8 words per thread), so the practical read + write ceiling that the
// HBM-bound stages are measured against.
typedef unsigned int copy_word_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const copy_word_t *__restrict__ src, copy_word_t *__restrict__ dst,
                                                size_t words) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

// ---- scratch of the allocation-free entry point (oct_radius_lists_noalloc)
// The in-place exclusive scan of the per-query entry counts keeps one partial sum per block of
// kNaChunk queries in `scratch`, a file-scope device array: each GPU holds its own copy, whose address
// oct_create resolves once per device (hipGetSymbolAddress, that device current) into g_scratch. Every
// call on a GPU shares that GPU's array; calls that overlap on one GPU race on it.
constexpr uint32_t kNaChunk = 8192;                                          // queries per scan block
constexpr uint32_t kNaMaxBlocks = 8192;                                      // partial sums held
constexpr size_t kNoAllocMaxQueries = (size_t)kNaChunk * kNaMaxBlocks;       // 67,108,864 queries per call
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint64_t scratch[kNaMaxBlocks];
std::mutex g_scratch_mu;
uint64_t *g_scratch[kMaxDevices] = {};

// Resolves this device's scratch address (the device must be current). Idempotent.
hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipSuccess;  // such a device has no allocation-free entry
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch[device] = (uint64_t *)p;
    return e;
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x, int lane) {
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}
// Exclusive prefix over a block of 256 threads (4 waves): returns the thread's exclusive prefix and
// the block total.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, uint64_t &total) {
    __shared__ uint64_t wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t inc = wave_incl_scan(x, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t base = 0;
    for (int i = 0; i < w; i++) base += wsum[i];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return base + inc - x;
}
// pass 1: part[b] = sum of v[b * kNaChunk .. ) (32 consecutive values per thread)
__global__ __launch_bounds__(256) void k_na_sums(const uint64_t *__restrict__ v, size_t n, uint64_t *__restrict__ part) {
    const size_t lo = (size_t)blockIdx.x * kNaChunk + (size_t)threadIdx.x * 32;
    uint64_t x = 0;
    for (size_t i = lo; i < lo + 32 && i < n; i++) x += v[i];
    uint64_t total;
    (void)block_excl_scan(x, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}
// pass 2 (one block): part[0 .. nb) -> exclusive prefixes in place; the grand total to *total
__global__ __launch_bounds__(256) void k_na_parts(uint64_t *__restrict__ part, uint32_t nb, uint64_t *__restrict__ total) {
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
        const uint32_t b = b0 + threadIdx.x;
        const uint64_t x = b < nb ? part[b] : 0;
        uint64_t t;
        const uint64_t e = block_excl_scan(x, t);
        if (b < nb) part[b] = carry + e;
        carry += t;
    }
    if (threadIdx.x == 0) *total = carry;
}
// pass 3: each block's values -> exclusive prefixes in place, offset by its partial
__global__ __launch_bounds__(256) void k_na_scan(uint64_t *__restrict__ v, size_t n, const uint64_t *__restrict__ part) {
    const size_t lo = (size_t)blockIdx.x * kNaChunk + (size_t)threadIdx.x * 32;
    uint64_t x = 0;
    for (size_t i = lo; i < lo + 32 && i < n; i++) x += v[i];
    uint64_t total;
    uint64_t run = part[blockIdx.x] + block_excl_scan(x, total);
    for (size_t i = lo; i < lo + 32 && i < n; i++) {
        const uint64_t c = v[i];
        v[i] = run;
        run += c;
    }
}
__global__ void k_na_zero(uint64_t *p) { *p = 0; }

// Copies n points given as three host coordinate arrays into the packed device array.
static int upload_xyz_soa(oct_ctx *ctx, const double *x, const double *y, const double *z, size_t n) {
    if (n >= (1ull << 30)) return fail(ctx, OCT_ERR_ARG, "at most 2^30 - 1 points");
    ctx->las_loaded = false;
    HCHK(ctx->xyz.reserve(3 * n), "alloc points");
    if (n) {
        if (!x || !y || !z) return fail(ctx, OCT_ERR_ARG, "null coordinate array");
        HCHK(ctx->qxyz.reserve(3 * n), "alloc staging");
        HCHK(hipMemcpyAsync(ctx->qxyz.p, x, n * 8, hipMemcpyHostToDevice, ctx->stream), "copy x");
        HCHK(hipMemcpyAsync(ctx->qxyz.p + n, y, n * 8, hipMemcpyHostToDevice, ctx->stream), "copy y");
        HCHK(hipMemcpyAsync(ctx->qxyz.p + 2 * n, z, n * 8, hipMemcpyHostToDevice, ctx->stream), "copy z");
        unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        k_interleave<<<nb, 256, 0, ctx->stream>>>(ctx->qxyz.p, ctx->xyz.p, n);
        HCHK(hipGetLastError(), "interleave points");
    }
    ctx->n = n;
    return OCT_OK;
}

// Copies n points (x, y, z leading doubles, stride bytes apart) into the packed device array.
// Packed input is one flat copy; strided input (e.g. the 32-byte Point) is copied flat into a
// staging buffer and repacked on the device.
int upload_xyz(oct_ctx *ctx, const double *src, size_t n, size_t stride, hipMemcpyKind kind) {
    if (stride < 24 || stride % 8) return fail(ctx, OCT_ERR_ARG, "stride_bytes must be >= 24 and a multiple of 8");
    if (n >= (1ull << 30)) return fail(ctx, OCT_ERR_ARG, "at most 2^30 - 1 points");
    ctx->las_loaded = false;  // the cloud no longer comes from a LAS file
    HCHK(ctx->xyz.reserve(3 * n), "alloc points");
    if (n) {
        if (!src) return fail(ctx, OCT_ERR_ARG, "null point buffer");
        if (stride == 24) {
            HCHK(hipMemcpyAsync(ctx->xyz.p, src, n * 24, kind, ctx->stream), "copy points");
        } else {
            const size_t words = n * (stride / 8);
            HCHK(ctx->qxyz.reserve(words), "alloc staging");
            HCHK(hipMemcpyAsync(ctx->qxyz.p, src, n * stride, kind, ctx->stream), "copy points");
            unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
            k_unstride<<<nb, 256, 0, ctx->stream>>>(ctx->qxyz.p, stride / 8, ctx->xyz.p, n);
            HCHK(hipGetLastError(), "repack points");
        }
    }
    ctx->n = n;
    return OCT_OK;
}

// Box(center, radii): min' = center - radii, max' = center + radii (box.hpp:17); halfLengths and
// precomputedRadii as setupBbox (linear_octree.hpp:439-455).
void setup_geometry(oct_ctx *ctx) {
    const double eps = (double)(1.0f / (float)(1 << kMaxDepth));
    for (int a = 0; a < 3; a++) {
        double mn = ctx->center[a] - ctx->radii[a], mx = ctx->center[a] + ctx->radii[a];
        ctx->boxmin[a] = mn;
        ctx->half_lengths[a] = (double)0.5f * eps * (mx - mn);
    }
    for (uint32_t l = 0; l <= kMaxDepth; l++) {
        double side = (double)(1u << (kMaxDepth - l));
        for (int a = 0; a < 3; a++) ctx->level_radii[l][a] = side * ctx->half_lengths[a];
    }
}

double stage_ms(const oct_ctx *ctx, Stage s) {
    if (!ctx->ev_used[s]) return 0.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0[s], ctx->ev1[s]) != hipSuccess) return 0.0;
    return (double)ms;
}

// Slot order that visits queries in SFC order (coherent waves): indices into the sorted cloud are
// already SFC positions; explicit centres are keyed with the cloud's encoder and box.
void query_order(oct_ctx *ctx, const double *centers, const uint32_t *idx, size_t nq, std::vector<uint32_t> &order) {
    order.resize(nq);
    std::iota(order.begin(), order.end(), 0u);
    if (idx) {
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return idx[a] < idx[b]; });
    } else if (centers) {
        std::vector<uint64_t> k(nq);
        for (size_t i = 0; i < nq; i++) {
            uint32_t x = quantize_axis(centers[3 * i], ctx->center[0], ctx->radii[0]);
            uint32_t y = quantize_axis(centers[3 * i + 1], ctx->center[1], ctx->radii[1]);
            uint32_t z = quantize_axis(centers[3 * i + 2], ctx->center[2], ctx->radii[2]);
            k[i] = ctx->encoder == OCT_ENC_HILBERT3D ? hilbert_encode_tab(kHilbert.t, x, y, z) : morton_encode(x, y, z);
        }
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    }
}

}  // namespace

// Upload of an already sorted cloud (points through `upload`) with its keys and Box.
template <class Upload>
static int set_sorted_impl(oct_ctx *ctx, size_t n, const uint64_t *codes, int encoder, const double center[3],
                           const double radii[3], Upload &&upload) {
    if (!ctx || !center || !radii || (n && !codes)) return fail(ctx, OCT_ERR_ARG, "null argument");
    if (!valid_encoder(encoder)) return fail(ctx, OCT_ERR_UNSUPPORTED, "only the 3-D encoders build octrees");
    HCHK(hipSetDevice(ctx->device), "set device");
    for (size_t i = 1; i < n; i++)
        if (codes[i] < codes[i - 1]) return fail(ctx, OCT_ERR_ARG, "codes must be sorted ascending");
    int r = upload();
    if (r) return r;
    HCHK(ctx->keys[0].reserve(n), "alloc keys");
    HCHK(ctx->p64.reserve(n), "alloc points");
    if (n) {
        HCHK(hipMemcpyAsync(ctx->keys[0].p, codes, n * 8, hipMemcpyHostToDevice, ctx->stream), "copy codes");
        unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        k_pad<<<nb, 256, 0, ctx->stream>>>(ctx->xyz.p, ctx->p64.p, n);
        HCHK(hipGetLastError(), "pad points");
    }
    HCHK(hipStreamSynchronize(ctx->stream), "set sorted");
    ctx->sorted_keys = ctx->keys[0].p;
    ctx->perm = nullptr;
    ctx->encoder = encoder;
    for (int a = 0; a < 3; a++) ctx->center[a] = center[a], ctx->radii[a] = radii[a];
    ctx->has_points = true;
    ctx->sort_lsd_learned = false;
    ctx->has_keys = false;
    ctx->has_sorted = true;
    ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

extern "C" {

const char *oct_version(void) { return "octree-amd 0.1 (gfx950)"; }

int oct_create(oct_ctx **out, int device) {
    if (!out) return OCT_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OCT_ERR_HIP;
    if (device < 0 || device >= ndev) return OCT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return OCT_ERR_HIP;
    oct_ctx *ctx = new oct_ctx;
    ctx->device = device;
    for (auto &a : ctx->group_tuned)
        for (auto &b : a)
            for (auto &c : b) c = -1;
    for (auto &a : ctx->knn_lanes_tuned)
        for (auto &b : a) b = -1;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return OCT_ERR_HIP;
    }
    for (int s = 0; s < ST_N; s++) {
        if (hipEventCreate(&ctx->ev0[s]) != hipSuccess || hipEventCreate(&ctx->ev1[s]) != hipSuccess) {
            delete ctx;
            return OCT_ERR_HIP;
        }
    }
    if (resolve_scratch(device) != hipSuccess) {
        delete ctx;
        return OCT_ERR_HIP;
    }
    if (ctx->derr.reserve(4) != hipSuccess || hipMemset(ctx->derr.p, 0, 4 * sizeof(uint32_t)) != hipSuccess) {
        ctx->derr.release();
        delete ctx;
        return OCT_ERR_HIP;
    }
    *out = ctx;
    return OCT_OK;
}

int oct_destroy(oct_ctx *ctx) {
    if (!ctx) return OCT_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->xyz.release();
    for (auto &b : ctx->keys) b.release();
    for (auto &b : ctx->vals) b.release();
    ctx->hist.release(), ctx->lookback.release(), ctx->misc.release();
    ctx->p64.release(), ctx->p32.release();
    ctx->nodes.release(), ctx->child.release(), ctx->nkey.release(), ctx->flags.release(), ctx->nrank.release();
    ctx->amax.release(), ctx->scan.release();
    ctx->qidx.release(), ctx->qxyz.release(), ctx->qrad.release(), ctx->counts.release(), ctx->qoff.release();
    ctx->entries.release(), ctx->stats.release(), ctx->kidx.release(), ctx->kd2.release();
    ctx->next.release(), ctx->qrad3.release(), ctx->kc_xyz.release(), ctx->kc_keys.release();
    ctx->kd2f.release(), ctx->radtab.release(), ctx->las_rec.release(), ctx->las_meta.release();
    ctx->rec_in.release(), ctx->rec_out.release(), ctx->derr.release();
    for (int s = 0; s < ST_N; s++) {
        if (ctx->ev0[s]) (void)hipEventDestroy(ctx->ev0[s]);
        if (ctx->ev1[s]) (void)hipEventDestroy(ctx->ev1[s]);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OCT_OK;
}

const char *oct_last_error(const oct_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int oct_synchronize(oct_ctx *ctx) {
    if (!ctx) return OCT_ERR_ARG;
    HCHK(hipStreamSynchronize(ctx->stream), "synchronize");
    return check_device_error(ctx);
}

int oct_set_points(oct_ctx *ctx, const double *xyz, size_t n, size_t stride_bytes) {
    if (!ctx) return OCT_ERR_ARG;
    HCHK(hipSetDevice(ctx->device), "set device");
    int r = upload_xyz(ctx, xyz, n, stride_bytes, hipMemcpyHostToDevice);
    if (r) return r;
    HCHK(hipStreamSynchronize(ctx->stream), "copy points");
    ctx->has_points = true;
    ctx->sort_lsd_learned = false;
    ctx->has_keys = ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

int oct_set_points_soa(oct_ctx *ctx, const double *x, const double *y, const double *z, size_t n) {
    if (!ctx) return OCT_ERR_ARG;
    HCHK(hipSetDevice(ctx->device), "set device");
    int r = upload_xyz_soa(ctx, x, y, z, n);
    if (r) return r;
    HCHK(hipStreamSynchronize(ctx->stream), "copy points");
    ctx->has_points = true;
    ctx->sort_lsd_learned = false;
    ctx->has_keys = ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

int oct_set_points_dev(oct_ctx *ctx, const double *xyz_dev, size_t n, size_t stride_bytes) {
    if (!ctx) return OCT_ERR_ARG;
    HCHK(hipSetDevice(ctx->device), "set device");
    int r = upload_xyz(ctx, xyz_dev, n, stride_bytes, hipMemcpyDeviceToDevice);
    if (r) return r;
    ctx->has_points = true;
    ctx->sort_lsd_learned = false;
    ctx->has_keys = ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

int oct_bbox(oct_ctx *ctx, double center[3], double radii[3]) {
    if (!ctx || !center || !radii) return fail(ctx, OCT_ERR_ARG, "null argument");
    if (!ctx->has_points) return fail(ctx, OCT_ERR_STATE, "oct_bbox before oct_set_points");
    HCHK(hipSetDevice(ctx->device), "set device");
    double mn[3], mx[3];
    if (ctx->n == 0) {
        for (int a = 0; a < 3; a++) mn[a] = DBL_MAX, mx[a] = -DBL_MAX;
    } else {
        HCHK(launch_bbox(ctx, mn, mx), "bbox");
    }
    // mbb: center = midpoint(min, max) = min + (max - min) / 2 (util.hpp:139-143); radii = (max - min) / 2
    for (int a = 0; a < 3; a++) {
        center[a] = mn[a] + (mx[a] - mn[a]) / 2;
        radii[a] = (mx[a] - mn[a]) / 2;
    }
    return OCT_OK;
}

int oct_encode(oct_ctx *ctx, int encoder, const double center[3], const double radii[3], uint64_t *keys_out) {
    if (!ctx || !center || !radii) return fail(ctx, OCT_ERR_ARG, "null argument");
    if (!valid_encoder(encoder)) return fail(ctx, OCT_ERR_UNSUPPORTED, "only MortonEncoder3D (1) and HilbertEncoder3D (5) build octrees");
    if (!ctx->has_points) return fail(ctx, OCT_ERR_STATE, "oct_encode before oct_set_points");
    HCHK(hipSetDevice(ctx->device), "set device");
    HCHK(launch_encode(ctx, encoder, center, radii), "encode");
    ctx->encoder = encoder;
    for (int a = 0; a < 3; a++) ctx->center[a] = center[a], ctx->radii[a] = radii[a];
    if (keys_out && ctx->n) HCHK(hipMemcpyAsync(keys_out, ctx->keys[0].p, ctx->n * 8, hipMemcpyDeviceToHost, ctx->stream), "copy keys");
    HCHK(hipStreamSynchronize(ctx->stream), "encode");
    ctx->has_keys = true;
    ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

int oct_sort(oct_ctx *ctx, uint32_t *perm_out, uint64_t *keys_out) {
    if (!ctx) return OCT_ERR_ARG;
    if (!ctx->has_keys) return fail(ctx, OCT_ERR_STATE, "oct_sort before oct_encode");
    HCHK(hipSetDevice(ctx->device), "set device");
    std::string err;
    hipError_t e = launch_sort(ctx, err);
    if (e != hipSuccess) return err.empty() ? hfail(ctx, e, "sort") : fail(ctx, OCT_ERR_STATE, err);
    HCHK(launch_gather(ctx), "gather");
    if (perm_out && ctx->n) HCHK(hipMemcpyAsync(perm_out, ctx->perm, ctx->n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy perm");
    if (keys_out && ctx->n)
        HCHK(hipMemcpyAsync(keys_out, ctx->sorted_keys, ctx->n * 8, hipMemcpyDeviceToHost, ctx->stream), "copy keys");
    HCHK(hipStreamSynchronize(ctx->stream), "sort");
    ctx->has_sorted = true;
    ctx->has_keys = false;  // the unsorted key buffer was consumed
    ctx->has_tree = ctx->has_lists = false;
    return OCT_OK;
}

int oct_gather_records(oct_ctx *ctx, const void *in, void *out, size_t record_bytes, int device_pointers) {
    if (!ctx) return OCT_ERR_ARG;
    if (!ctx->has_sorted || !ctx->perm) return fail(ctx, OCT_ERR_STATE, "oct_gather_records needs the permutation of oct_sort");
    if (record_bytes == 0 || record_bytes > (1u << 16)) return fail(ctx, OCT_ERR_ARG, "record_bytes must be 1 .. 65536");
    const size_t n = ctx->n;
    if (!n) return OCT_OK;
    if (!in || !out) return fail(ctx, OCT_ERR_ARG, "null record buffer");
    HCHK(hipSetDevice(ctx->device), "set device");
    if (device_pointers) {
        HCHK(launch_gather_records(ctx->stream, in, ctx->perm, out, n, record_bytes), "gather records");
    } else {
        HCHK(ctx->rec_in.reserve(n * record_bytes), "alloc records");
        HCHK(ctx->rec_out.reserve(n * record_bytes), "alloc records");
        HCHK(hipMemcpyAsync(ctx->rec_in.p, in, n * record_bytes, hipMemcpyHostToDevice, ctx->stream), "copy records");
        HCHK(launch_gather_records(ctx->stream, ctx->rec_in.p, ctx->perm, ctx->rec_out.p, n, record_bytes), "gather records");
        HCHK(hipMemcpyAsync(out, ctx->rec_out.p, n * record_bytes, hipMemcpyDeviceToHost, ctx->stream), "copy records");
    }
    HCHK(hipStreamSynchronize(ctx->stream), "gather records");
    return OCT_OK;
}

// oct_bbox + oct_encode + oct_sort with nothing read back in between: the box stays on the device
// (k_bbox_final computes the centre and radii with the host's expressions), the encode reads it there,
// and the host reads it back after the sort.
int oct_sort_points(oct_ctx *ctx, int encoder, double center_out[3], double radii_out[3], uint32_t *perm_out,
                    uint64_t *keys_out) {
    if (!ctx) return OCT_ERR_ARG;
    if (!valid_encoder(encoder)) return fail(ctx, OCT_ERR_UNSUPPORTED, "only MortonEncoder3D (1) and HilbertEncoder3D (5) build octrees");
    if (!ctx->has_points) return fail(ctx, OCT_ERR_STATE, "oct_sort_points before oct_set_points");
    double c[3], r[3];
    int s;
    if (ctx->n == 0) {  // the empty box of oct_bbox, through the separate calls
        if ((s = oct_bbox(ctx, c, r))) return s;
        if ((s = oct_encode(ctx, encoder, c, r, nullptr))) return s;
        if ((s = oct_sort(ctx, perm_out, keys_out))) return s;
    } else {
        HCHK(hipSetDevice(ctx->device), "set device");
        HCHK(launch_bbox_dev(ctx), "bbox");
        HCHK(launch_encode(ctx, encoder, nullptr, nullptr), "encode");
        ctx->encoder = encoder;
        ctx->has_keys = true;
        ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
        s = oct_sort(ctx, perm_out, keys_out);  // ends with a stream synchronisation
        // the box belongs to the encoded keys, which stay valid for another oct_sort if this one failed
        double b[12];
        HCHK(hipMemcpyAsync(b, ctx->box.p, sizeof(b), hipMemcpyDeviceToHost, ctx->stream), "copy box");
        HCHK(hipStreamSynchronize(ctx->stream), "copy box");
        for (int a = 0; a < 3; a++) c[a] = b[6 + a], r[a] = b[9 + a];
        for (int a = 0; a < 3; a++) ctx->center[a] = c[a], ctx->radii[a] = r[a];
        if (s) return s;
    }
    if (center_out)
        for (int a = 0; a < 3; a++) center_out[a] = c[a];
    if (radii_out)
        for (int a = 0; a < 3; a++) radii_out[a] = r[a];
    return OCT_OK;
}

int oct_set_sorted(oct_ctx *ctx, const double *xyz, size_t n, size_t stride_bytes, const uint64_t *codes, int encoder,
                   const double center[3], const double radii[3]) {
    return set_sorted_impl(ctx, n, codes, encoder, center, radii,
                           [&] { return upload_xyz(ctx, xyz, n, stride_bytes, hipMemcpyHostToDevice); });
}

int oct_set_sorted_soa(oct_ctx *ctx, const double *x, const double *y, const double *z, size_t n, const uint64_t *codes,
                       int encoder, const double center[3], const double radii[3]) {
    return set_sorted_impl(ctx, n, codes, encoder, center, radii, [&] { return upload_xyz_soa(ctx, x, y, z, n); });
}

int oct_build(oct_ctx *ctx, uint32_t max_points_leaf, oct_tree_info *info) {
    if (!ctx) return OCT_ERR_ARG;
    if (!ctx->has_sorted) return fail(ctx, OCT_ERR_STATE, "oct_build before the cloud is sorted");
    HCHK(hipSetDevice(ctx->device), "set device");
    setup_geometry(ctx);
    std::string err;
    hipError_t e = launch_build(ctx, max_points_leaf, err);
    if (e != hipSuccess) return err.empty() ? hfail(ctx, e, "build") : fail(ctx, OCT_ERR_STATE, err);
    HCHK(hipStreamSynchronize(ctx->stream), "build");
    ctx->has_tree = true;
    ctx->has_lists = false;
    if (info) {
        info->n_leaf = ctx->n_leaf;
        info->n_internal = ctx->n_internal;
        info->n_total = ctx->n_total;
        info->max_depth_seen = ctx->max_depth_seen;
        info->max_points_leaf = max_points_leaf;
        info->n_levels = ctx->n_levels;
        info->build_ms = stage_ms(ctx, ST_BUILD);
    }
    return OCT_OK;
}

int oct_get_tree(oct_ctx *ctx, uint64_t *prefixes, uint32_t *offsets, uint64_t *ranges, double *centers, double *radii) {
    if (!ctx) return OCT_ERR_ARG;
    if (!ctx->has_tree) return fail(ctx, OCT_ERR_STATE, "oct_get_tree before oct_build");
    HCHK(hipSetDevice(ctx->device), "set device");
    const size_t nt = ctx->n_total;
    std::vector<NodeRec> nodes(nt);
    std::vector<uint64_t> keys(nt);
    HCHK(hipMemcpyAsync(nodes.data(), ctx->nodes.p, nt * sizeof(NodeRec), hipMemcpyDeviceToHost, ctx->stream), "copy nodes");
    HCHK(hipMemcpyAsync(keys.data(), ctx->nkey.p, nt * 8, hipMemcpyDeviceToHost, ctx->stream), "copy keys");
    if (offsets) HCHK(hipMemcpyAsync(offsets, ctx->child.p, nt * 4, hipMemcpyDeviceToHost, ctx->stream), "copy offsets");
    HCHK(hipStreamSynchronize(ctx->stream), "get tree");
    for (uint32_t l = 0; l <= kMaxDepth; l++)
        for (uint32_t u = ctx->level_start[l]; u < ctx->level_start[l + 1] && u < nt; u++) {
            if (prefixes) prefixes[u] = encode_placeholder(keys[u], l);
        }
    for (size_t u = 0; u < nt; u++) {
        if (ranges) ranges[2 * u] = nodes[u].first, ranges[2 * u + 1] = nodes[u].last;
        if (centers) centers[3 * u] = nodes[u].cx, centers[3 * u + 1] = nodes[u].cy, centers[3 * u + 2] = nodes[u].cz;
    }
    if (radii)
        for (int l = 0; l < 22; l++)
            for (int a = 0; a < 3; a++) radii[3 * l + a] = ctx->level_radii[l][a];
    return OCT_OK;
}

static int check_search_args(oct_ctx *ctx, int kernel, int algo, int precision, bool approx = false) {
    if (kernel < OCT_KERNEL_CIRCLE || kernel > OCT_KERNEL_CUBE) return fail(ctx, OCT_ERR_ARG, "unknown kernel id");
    if (algo == OCT_ALGO_APPROX && !approx) return fail(ctx, OCT_ERR_ARG, "neighborsApprox runs through oct_radius_approx");
    if (algo < OCT_ALGO_NEIGHBORS || algo > OCT_ALGO_APPROX) return fail(ctx, OCT_ERR_ARG, "unknown search algorithm id");
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return fail(ctx, OCT_ERR_ARG, "unknown precision");
    if (!ctx->has_tree) return fail(ctx, OCT_ERR_STATE, "search before oct_build");
    return OCT_OK;
}

// Queries per wave for the count-mode kernel: explicit override, else autotuned, else 32.
static int radius_group(const oct_ctx *ctx, int kernel, int algo, int precision) {
    if (ctx->group_override >= 0) return ctx->group_override;
    int t = ctx->group_tuned[kernel & 3][algo & 3][precision & 1];
    return t >= 0 ? t : 32;
}

static int read_stats(oct_ctx *ctx, oct_search_stats *stats, double ms) {
    unsigned long long h[6];
    HCHK(hipMemcpyAsync(h, ctx->stats.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "copy stats");
    HCHK(hipStreamSynchronize(ctx->stream), "stats");
    stats->nodes_classified = h[0];
    stats->points_tested = h[1];
    stats->rechecked = h[2];
    stats->total_neighbors = h[3];
    stats->kernel_ms = ms;
    stats->class_slots = h[4];
    stats->test_slots = h[5];
    return OCT_OK;
}

static int radius_impl(oct_ctx *ctx, int kernel, int algo, int precision, const double *centers,
                       const uint32_t *center_idx, const double *radii, double radius, const double *radii3,
                       const double *radius3, size_t nq, int want_lists, uint32_t *counts_out, oct_search_stats *stats,
                       uint32_t approx_level = ~0u, bool approx_upper = false) {
    int s = check_search_args(ctx, kernel, algo, precision, algo == OCT_ALGO_APPROX);
    if (s) return s;
    if (nq >= (1ull << 31)) return fail(ctx, OCT_ERR_ARG, "too many queries");
    if (!centers && !center_idx && nq > ctx->n) return fail(ctx, OCT_ERR_ARG, "all-points search with nq > n");
    if (center_idx)
        for (size_t i = 0; i < nq; i++)
            if (center_idx[i] >= ctx->n) return fail(ctx, OCT_ERR_ARG, "center index out of range");
    HCHK(hipSetDevice(ctx->device), "set device");
    std::vector<uint32_t> order;
    query_order(ctx, centers, center_idx, nq, order);
    HCHK(ctx->qidx.reserve(2 * nq + 1), "alloc queries");
    uint32_t *d_order = ctx->qidx.p, *d_idx = ctx->qidx.p + nq;
    if (nq) HCHK(hipMemcpyAsync(d_order, order.data(), nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy order");
    if (center_idx && nq) HCHK(hipMemcpyAsync(d_idx, center_idx, nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy idx");
    if (centers) {
        HCHK(ctx->qxyz.reserve(3 * nq + 1), "alloc centres");
        if (nq) HCHK(hipMemcpyAsync(ctx->qxyz.p, centers, nq * 24, hipMemcpyHostToDevice, ctx->stream), "copy centres");
    }
    if (radii) {
        HCHK(ctx->qrad.reserve(nq + 1), "alloc radii");
        if (nq) HCHK(hipMemcpyAsync(ctx->qrad.p, radii, nq * 8, hipMemcpyHostToDevice, ctx->stream), "copy radii");
    }
    if (radii3) {
        HCHK(ctx->qrad3.reserve(3 * nq + 1), "alloc radii");
        if (nq) HCHK(hipMemcpyAsync(ctx->qrad3.p, radii3, nq * 24, hipMemcpyHostToDevice, ctx->stream), "copy radii");
    }
    HCHK(ctx->counts.reserve(2 * nq + 1), "alloc counts");
    HCHK(ctx->stats.reserve(8), "alloc stats");
    HCHK(hipMemsetAsync(ctx->stats.p, 0, 8 * sizeof(unsigned long long), ctx->stream), "clear stats");
    RadiusLaunch L;
    L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = want_lists ? 1 : 0;
    L.group = radius_group(ctx, kernel, algo, precision);
    L.remap = ctx->remap_override != 0;
    L.approx_level = approx_level, L.approx_upper = approx_upper;
    L.order = d_order;
    L.qidx = center_idx ? d_idx : nullptr;
    L.qxyz = centers ? ctx->qxyz.p : nullptr;
    L.qrad = radii ? ctx->qrad.p : nullptr;
    L.radius = radius;
    if (radii3 || radius3) {
        L.use_r3 = true;
        L.qrad3 = radii3 ? ctx->qrad3.p : nullptr;
        if (radius3)
            for (int k = 0; k < 3; k++) L.radius3[k] = radius3[k];
    }
    L.nq = nq;
    L.counts = ctx->counts.p;
    L.ecounts = ctx->counts.p + nq;
    L.stats = ctx->stats.p;
    HCHK(launch_radius(ctx, L), "radius count");
    double ms = 0.0;
    HCHK(hipStreamSynchronize(ctx->stream), "radius count");
    ms += stage_ms(ctx, ST_SEARCH);
    ctx->has_lists = false;
    if (want_lists) {
        HCHK(ctx->qoff.reserve(nq + 1), "alloc offsets");
        HCHK(exclusive_scan_u32_to_u64(ctx, ctx->counts.p + nq, ctx->qoff.p, nq), "scan");
        uint64_t total = 0;
        HCHK(hipMemcpyAsync(&total, ctx->qoff.p + nq, 8, hipMemcpyDeviceToHost, ctx->stream), "copy total");
        HCHK(hipStreamSynchronize(ctx->stream), "scan");
        const size_t words = (size_t)total * (algo == OCT_ALGO_STRUCT || algo == OCT_ALGO_APPROX ? 2 : 1);
        HCHK(ctx->entries.reserve(words + 1), "alloc lists");
        L.mode = 2;
        L.offsets = ctx->qoff.p;
        L.entries = ctx->entries.p;
        L.stats = nullptr;
        HCHK(launch_radius(ctx, L), "radius fill");
        HCHK(hipStreamSynchronize(ctx->stream), "radius fill");
        ms += stage_ms(ctx, ST_SEARCH);
        ctx->has_lists = true;
    }
    ctx->last_nq = nq;
    ctx->last_algo = algo;
    if (counts_out && nq) HCHK(hipMemcpyAsync(counts_out, ctx->counts.p, nq * 4, hipMemcpyDeviceToHost, ctx->stream), "copy counts");
    HCHK(hipStreamSynchronize(ctx->stream), "radius");
    if ((s = check_device_error(ctx))) return s;
    if (stats) return read_stats(ctx, stats, ms);
    return OCT_OK;
}

int oct_radius(oct_ctx *ctx, int kernel, int algo, int precision, const double *centers, const uint32_t *center_idx,
               const double *radii, double radius, size_t nq, int want_lists, uint32_t *counts_out,
               oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    if (algo == OCT_ALGO_APPROX) return fail(ctx, OCT_ERR_ARG, "neighborsApprox runs through oct_radius_approx");
    return radius_impl(ctx, kernel, algo, precision, centers, center_idx, radii, radius, nullptr, nullptr, nq, want_lists,
                       counts_out, stats);
}

int oct_radius_lists_noalloc(oct_ctx *ctx, int kernel, int algo, int precision, const uint32_t *center_idx_dev,
                             double radius, size_t nq, uint32_t *counts_dev, uint64_t *offsets_dev, uint32_t *entries_dev,
                             size_t entries_capacity, uint64_t *total_out) {
    if (!ctx) return OCT_ERR_ARG;
    int s = check_search_args(ctx, kernel, algo, precision);
    if (s) return s;
    if (!counts_dev || !offsets_dev || !total_out) return fail(ctx, OCT_ERR_ARG, "counts, offsets and total are required");
    if (nq > kNoAllocMaxQueries)
        return fail(ctx, OCT_ERR_ARG, "at most " + std::to_string(kNoAllocMaxQueries) + " queries per allocation-free call");
    if (!center_idx_dev && nq > ctx->n) return fail(ctx, OCT_ERR_ARG, "all-points search with nq > n");
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(ctx, OCT_ERR_STATE, "no scratch resolved for this device");
    uint64_t *part = g_scratch[ctx->device];
    HCHK(hipSetDevice(ctx->device), "set device");
    *total_out = 0;
    // sizes: per-query neighbour counts and entry counts (indices, or ranges for Struct)
    RadiusLaunch L;
    L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 1;
    L.qidx = center_idx_dev;
    L.radius = radius;
    L.nq = nq;
    L.counts = counts_dev;
    L.ecounts64 = offsets_dev;
    HCHK(launch_radius(ctx, L), "radius sizes");
    // offsets: exclusive prefix of the entry counts, in place; offsets[nq] = total
    if (nq) {
        const uint32_t nb = (uint32_t)((nq + kNaChunk - 1) / kNaChunk);
        k_na_sums<<<nb, 256, 0, ctx->stream>>>(offsets_dev, nq, part);
        k_na_parts<<<1, 256, 0, ctx->stream>>>(part, nb, offsets_dev + nq);
        k_na_scan<<<nb, 256, 0, ctx->stream>>>(offsets_dev, nq, part);
    } else {
        k_na_zero<<<1, 1, 0, ctx->stream>>>(offsets_dev);
    }
    HCHK(hipGetLastError(), "offsets");
    uint64_t total = 0;
    HCHK(hipMemcpyAsync(&total, offsets_dev + nq, 8, hipMemcpyDeviceToHost, ctx->stream), "copy total");
    HCHK(hipStreamSynchronize(ctx->stream), "radius sizes");
    if ((s = check_device_error(ctx))) return s;
    *total_out = total;
    const size_t words = (size_t)total * (algo == OCT_ALGO_STRUCT ? 2 : 1);
    if (words > entries_capacity)
        return fail(ctx, OCT_ERR_ARG, "entries_capacity " + std::to_string(entries_capacity) + " is below the " +
                                          std::to_string(words) + " words the lists need (counts and offsets are valid)");
    if (words && !entries_dev) return fail(ctx, OCT_ERR_ARG, "null entries buffer");
    // fill: every query's list at its offset
    L.mode = 2;
    L.counts = nullptr;
    L.ecounts64 = nullptr;
    L.offsets = offsets_dev;
    L.entries = entries_dev;
    if (words) HCHK(launch_radius(ctx, L), "radius fill");
    HCHK(hipStreamSynchronize(ctx->stream), "radius fill");
    ctx->has_lists = false;
    return check_device_error(ctx);
}

int oct_radius_box(oct_ctx *ctx, int kernel, int algo, const double *centers, const uint32_t *center_idx,
                   const double *radii3, const double radius3[3], size_t nq, int want_lists, uint32_t *counts_out,
                   oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    if (algo == OCT_ALGO_APPROX) return fail(ctx, OCT_ERR_ARG, "neighborsApprox runs through oct_radius_approx");
    if (kernel != OCT_KERNEL_CUBE && kernel != OCT_KERNEL_SQUARE)
        return fail(ctx, OCT_ERR_UNSUPPORTED, "per-axis radii need the cube or square kernel (kernel_factory.hpp:33-41)");
    if (!radii3 && !radius3) return fail(ctx, OCT_ERR_ARG, "no radii");
    return radius_impl(ctx, kernel, algo, OCT_PREC_FP64, centers, center_idx, nullptr, 0.0, radii3, radius3, nq,
                       want_lists, counts_out, stats);
}

int oct_radius_approx(oct_ctx *ctx, int kernel, int precision, const double *centers, const uint32_t *center_idx,
                      const double *radii, double radius, const double radius3[3], size_t nq, uint32_t level, int upper_bound,
                      int want_lists, uint32_t *counts_out, oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    if (radius3 && kernel != OCT_KERNEL_CUBE && kernel != OCT_KERNEL_SQUARE)
        return fail(ctx, OCT_ERR_UNSUPPORTED, "per-axis radii need the cube or square kernel (kernel_factory.hpp:33-41)");
    return radius_impl(ctx, kernel, OCT_ALGO_APPROX, radius3 ? OCT_PREC_FP64 : precision, centers, center_idx, radii, radius,
                       nullptr, radius3, nq, want_lists, counts_out, stats, level, upper_bound != 0);
}

static int key_codec(oct_ctx *ctx, int encoder, bool decode, uint32_t *xyz, uint64_t *keys, size_t n) {
    if (!ctx) return OCT_ERR_ARG;
    if (encoder != OCT_ENC_MORTON3D && encoder != OCT_ENC_HILBERT3D) return fail(ctx, OCT_ERR_UNSUPPORTED, "encoder");
    if (n && (!xyz || !keys)) return fail(ctx, OCT_ERR_ARG, "null argument");
    if (!n) return OCT_OK;
    HCHK(hipSetDevice(ctx->device), "set device");
    HCHK(ctx->kc_xyz.reserve(3 * n), "alloc coordinates");
    HCHK(ctx->kc_keys.reserve(n), "alloc keys");
    if (decode)
        HCHK(hipMemcpyAsync(ctx->kc_keys.p, keys, n * 8, hipMemcpyHostToDevice, ctx->stream), "copy keys");
    else
        HCHK(hipMemcpyAsync(ctx->kc_xyz.p, xyz, n * 12, hipMemcpyHostToDevice, ctx->stream), "copy coordinates");
    HCHK(launch_key_codec(ctx, encoder, decode, ctx->kc_xyz.p, ctx->kc_keys.p, n), "key codec");
    if (decode)
        HCHK(hipMemcpyAsync(xyz, ctx->kc_xyz.p, n * 12, hipMemcpyDeviceToHost, ctx->stream), "copy coordinates");
    else
        HCHK(hipMemcpyAsync(keys, ctx->kc_keys.p, n * 8, hipMemcpyDeviceToHost, ctx->stream), "copy keys");
    HCHK(hipStreamSynchronize(ctx->stream), "key codec");
    return OCT_OK;
}

int oct_encode_coords(oct_ctx *ctx, int encoder, const uint32_t *xyz, size_t n, uint64_t *keys) {
    return key_codec(ctx, encoder, false, const_cast<uint32_t *>(xyz), keys, n);
}

int oct_decode_keys(oct_ctx *ctx, int encoder, const uint64_t *keys, size_t n, uint32_t *xyz) {
    return key_codec(ctx, encoder, true, xyz, const_cast<uint64_t *>(keys), n);
}

int oct_radius_fetch(oct_ctx *ctx, uint64_t *offsets, uint32_t *entries, size_t entries_capacity) {
    if (!ctx || !offsets) return fail(ctx, OCT_ERR_ARG, "null argument");
    if (!ctx->has_lists) return fail(ctx, OCT_ERR_STATE, "no result lists (call oct_radius with want_lists = 1)");
    HCHK(hipSetDevice(ctx->device), "set device");
    const size_t nq = ctx->last_nq;
    HCHK(hipMemcpyAsync(offsets, ctx->qoff.p, (nq + 1) * 8, hipMemcpyDeviceToHost, ctx->stream), "copy offsets");
    HCHK(hipStreamSynchronize(ctx->stream), "fetch");
    const size_t words = (size_t)offsets[nq] * (ctx->last_algo == OCT_ALGO_STRUCT || ctx->last_algo == OCT_ALGO_APPROX ? 2 : 1);
    if (words && entries) {  // entries == NULL: size query (offsets only)
        if (entries_capacity < words) return fail(ctx, OCT_ERR_ARG, "entries buffer too small");
        HCHK(hipMemcpyAsync(entries, ctx->entries.p, words * 4, hipMemcpyDeviceToHost, ctx->stream), "copy lists");
        HCHK(hipStreamSynchronize(ctx->stream), "fetch");
    }
    return OCT_OK;
}

int oct_radius_dev(oct_ctx *ctx, int kernel, int algo, int precision, const uint32_t *center_idx_dev, double radius,
                   size_t nq, uint32_t *counts_dev, oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    int s = check_search_args(ctx, kernel, algo, precision);
    if (s) return s;
    if (!center_idx_dev && nq > ctx->n) return fail(ctx, OCT_ERR_ARG, "all-points search with nq > n");
    HCHK(hipSetDevice(ctx->device), "set device");
    if (!counts_dev) {
        HCHK(ctx->counts.reserve(2 * nq + 1), "alloc counts");
        counts_dev = ctx->counts.p;
    }
    RadiusLaunch L;
    L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 0;
    L.group = radius_group(ctx, kernel, algo, precision);
    L.remap = ctx->remap_override != 0;
    L.qidx = center_idx_dev;
    L.radius = radius;
    L.nq = nq;
    L.counts = counts_dev;
    if (stats) {
        HCHK(ctx->stats.reserve(8), "alloc stats");
        HCHK(hipMemsetAsync(ctx->stats.p, 0, 8 * sizeof(unsigned long long), ctx->stream), "clear stats");
        L.stats = ctx->stats.p;
    }
    HCHK(launch_radius(ctx, L), "radius");
    ctx->has_lists = false;
    if (stats) {
        HCHK(hipStreamSynchronize(ctx->stream), "radius");
        if ((s = check_device_error(ctx))) return s;
        return read_stats(ctx, stats, stage_ms(ctx, ST_SEARCH));
    }
    return OCT_OK;
}

int oct_set_tuning(oct_ctx *ctx, int knob, int value) {
    if (!ctx) return OCT_ERR_ARG;
    if (knob == OCT_TUNE_RADIUS_MFMA) {  // the point tests run on the vector ALUs only
        if (value == 1)
            return fail(ctx, OCT_ERR_UNSUPPORTED, "no matrix-core point tests: the vector ALUs measured faster on every workload");
        if (value != -1 && value != 0) return fail(ctx, OCT_ERR_ARG, "radius mfma switch must be -1 or 0");
        return OCT_OK;
    }
    if (knob == OCT_TUNE_RADIUS_REMAP) {
        if (value != -1 && value != 0 && value != 1) return fail(ctx, OCT_ERR_ARG, "radius remap switch must be -1, 0 or 1");
        ctx->remap_override = value;
        return OCT_OK;
    }
    if (knob == OCT_TUNE_SORT) {
        if (value < -1 || value > 2) return fail(ctx, OCT_ERR_ARG, "sort switch must be -1, 0, 1 or 2");
        ctx->sort_override = value;
        return OCT_OK;
    }
    if (knob == OCT_TUNE_KNN_LANES) {
        if (value != -1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32 && value != 64)
            return fail(ctx, OCT_ERR_ARG, "knn lanes per query must be -1, 2, 4, 8, 16, 32 or 64");
        ctx->knn_lanes_override = value;
        return OCT_OK;
    }
    if (knob != OCT_TUNE_RADIUS_GROUP) return fail(ctx, OCT_ERR_ARG, "unknown tuning knob");
    if (value != -1 && value != 0 && value != 16 && value != 32 && value != 64)
        return fail(ctx, OCT_ERR_ARG, "radius group must be -1, 0, 16, 32 or 64");
    ctx->group_override = value;
    return OCT_OK;
}

int oct_autotune_radius(oct_ctx *ctx, int kernel, int algo, int precision, double radius, size_t sample_nq, int *best,
                        double ms_out[4]) {
    if (!ctx) return OCT_ERR_ARG;
    int s = check_search_args(ctx, kernel, algo, precision);
    if (s) return s;
    if (ctx->n == 0) return fail(ctx, OCT_ERR_STATE, "autotune on an empty cloud");
    HCHK(hipSetDevice(ctx->device), "set device");
    size_t nq = std::min(sample_nq ? sample_nq : (size_t)1 << 20, ctx->n);
    // 16 chunks of consecutive SFC positions spread over the cloud
    std::vector<uint32_t> idx(nq);
    const size_t chunks = 16, per = (nq + chunks - 1) / chunks;
    size_t k = 0;
    for (size_t c = 0; c < chunks && k < nq; c++) {
        size_t start = (ctx->n / chunks) * c;
        for (size_t i = 0; i < per && k < nq; i++) idx[k++] = (uint32_t)std::min(start + i, ctx->n - 1);
    }
    HCHK(ctx->qidx.reserve(nq + 1), "alloc queries");
    HCHK(ctx->counts.reserve(nq + 1), "alloc counts");
    HCHK(hipMemcpyAsync(ctx->qidx.p, idx.data(), nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy sample");
    const int cand[4] = {0, 16, 32, 64};
    double bestms = 1e300;
    int bestg = 32;
    for (int ci = 0; ci < 4; ci++) {
        RadiusLaunch L;
        L.kernel = kernel, L.algo = algo, L.precision = precision, L.mode = 0, L.group = cand[ci];
        L.remap = ctx->remap_override != 0;
        L.qidx = ctx->qidx.p;
        L.radius = radius;
        L.nq = nq;
        L.counts = ctx->counts.p;
        double t = 1e300;
        for (int rep = 0; rep < 3; rep++) {  // first repetition warms caches and code objects
            HCHK(launch_radius(ctx, L), "autotune");
            HCHK(hipStreamSynchronize(ctx->stream), "autotune");
            if (rep) t = std::min(t, stage_ms(ctx, ST_SEARCH));
        }
        if (ms_out) ms_out[ci] = t;
        if (t < bestms) bestms = t, bestg = cand[ci];
    }
    ctx->group_tuned[kernel & 3][algo & 3][precision & 1] = bestg;
    ctx->has_lists = false;
    if (best) *best = bestg;
    return OCT_OK;
}

int oct_tuned_radius(oct_ctx *ctx, int kernel, int algo, int precision, int *group, int *mfma) {
    if (!ctx) return OCT_ERR_ARG;
    if (kernel < 0 || kernel > 3 || algo < 1 || algo > 3 || (precision != 0 && precision != 1))
        return fail(ctx, OCT_ERR_ARG, "unknown kernel, algorithm or precision");
    if (group) *group = radius_group(ctx, kernel, algo, precision);
    if (mfma) *mfma = 0;  // point tests on the vector ALUs
    return OCT_OK;
}

// kNN size class of k: 0 for k <= 16, 1 for <= 32, ... 6 for <= 1024
static int knn_class(uint32_t k) {
    int c = 0;
    while (c < 6 && (16u << c) < k) c++;
    return c;
}
// lanes per query for a kNN launch: explicit override, else autotuned, else the size class default
static int knn_lanes(const oct_ctx *ctx, int precision, uint32_t k) {
    if (ctx->knn_lanes_override > 0) return ctx->knn_lanes_override;
    const int t = ctx->knn_lanes_tuned[precision & 1][knn_class(k)];
    return t > 0 ? t : 0;
}

int oct_knn(oct_ctx *ctx, int precision, const double *centers, const uint32_t *center_idx, size_t nq, uint32_t k,
            int32_t *idx_out, double *d2_out, oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return fail(ctx, OCT_ERR_ARG, "unknown precision");
    if (!ctx->has_tree) return fail(ctx, OCT_ERR_STATE, "knn before oct_build");
    if (k == 0) return OCT_OK;
    if (center_idx)
        for (size_t i = 0; i < nq; i++)
            if (center_idx[i] >= ctx->n) return fail(ctx, OCT_ERR_ARG, "center index out of range");
    if (!centers && !center_idx && nq > ctx->n) return fail(ctx, OCT_ERR_ARG, "all-points kNN with nq > n");
    HCHK(hipSetDevice(ctx->device), "set device");
    std::vector<uint32_t> order;
    query_order(ctx, centers, center_idx, nq, order);
    HCHK(ctx->qidx.reserve(2 * nq + 1), "alloc queries");
    uint32_t *d_order = ctx->qidx.p, *d_idx = ctx->qidx.p + nq;
    if (nq) HCHK(hipMemcpyAsync(d_order, order.data(), nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy order");
    if (center_idx && nq) HCHK(hipMemcpyAsync(d_idx, center_idx, nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy idx");
    if (centers) {
        HCHK(ctx->qxyz.reserve(3 * nq + 1), "alloc centres");
        if (nq) HCHK(hipMemcpyAsync(ctx->qxyz.p, centers, nq * 24, hipMemcpyHostToDevice, ctx->stream), "copy centres");
    }
    HCHK(ctx->kidx.reserve((size_t)nq * k + 1), "alloc knn");
    HCHK(ctx->kd2.reserve((size_t)nq * k + 1), "alloc knn");
    HCHK(ctx->stats.reserve(8), "alloc stats");
    HCHK(hipMemsetAsync(ctx->stats.p, 0, 8 * sizeof(unsigned long long), ctx->stream), "clear stats");
    KnnLaunch L;
    L.precision = precision;
    L.order = d_order;
    L.qidx = center_idx ? d_idx : nullptr;
    L.qxyz = centers ? ctx->qxyz.p : nullptr;
    L.nq = nq, L.k = k;
    L.idx = ctx->kidx.p, L.d2 = ctx->kd2.p;
    L.stats = ctx->stats.p;
    L.lanes = knn_lanes(ctx, precision, k);
    std::string err;
    hipError_t e = launch_knn(ctx, L, err);
    if (e != hipSuccess) return err.empty() ? hfail(ctx, e, "knn") : fail(ctx, OCT_ERR_ARG, err);
    if (idx_out && nq) HCHK(hipMemcpyAsync(idx_out, ctx->kidx.p, nq * k * 4, hipMemcpyDeviceToHost, ctx->stream), "copy idx");
    if (d2_out && nq) HCHK(hipMemcpyAsync(d2_out, ctx->kd2.p, nq * k * 8, hipMemcpyDeviceToHost, ctx->stream), "copy d2");
    HCHK(hipStreamSynchronize(ctx->stream), "knn");
    if (int de = check_device_error(ctx)) return de;
    if (stats) return read_stats(ctx, stats, stage_ms(ctx, ST_KNN));
    return OCT_OK;
}

int oct_knn_dev(oct_ctx *ctx, int precision, const uint32_t *center_idx_dev, size_t first, size_t nq, uint32_t k,
                int32_t *idx_dev, float *d2_dev, oct_search_stats *stats) {
    if (!ctx) return OCT_ERR_ARG;
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return fail(ctx, OCT_ERR_ARG, "unknown precision");
    if (!ctx->has_tree) return fail(ctx, OCT_ERR_STATE, "knn before oct_build");
    if (k > 1024) return fail(ctx, OCT_ERR_ARG, "knn: k above 1024 is not supported");
    if (!center_idx_dev && (first > ctx->n || nq > ctx->n - first))
        return fail(ctx, OCT_ERR_ARG, "all-points kNN range beyond the cloud");
    if (center_idx_dev && first) return fail(ctx, OCT_ERR_ARG, "first applies to all-points kNN only");
    HCHK(hipSetDevice(ctx->device), "set device");
    const bool keep = !idx_dev || !d2_dev;
    if (keep) {
        HCHK(ctx->kidx.reserve((size_t)nq * k + 1), "alloc knn");
        HCHK(ctx->kd2f.reserve((size_t)nq * k + 1), "alloc knn");
        ctx->knn_nq = 0, ctx->knn_k = 0;
    }
    KnnLaunch L;
    L.precision = precision;
    L.qidx = center_idx_dev;
    L.qoff = (uint32_t)first;
    L.nq = nq, L.k = k;
    L.idx = idx_dev ? idx_dev : ctx->kidx.p;
    L.d2f = d2_dev ? d2_dev : ctx->kd2f.p;
    L.lanes = knn_lanes(ctx, precision, k);
    if (stats) {
        HCHK(ctx->stats.reserve(8), "alloc stats");
        HCHK(hipMemsetAsync(ctx->stats.p, 0, 8 * sizeof(unsigned long long), ctx->stream), "clear stats");
        L.stats = ctx->stats.p;
    }
    std::string err;
    hipError_t e = launch_knn(ctx, L, err);
    if (e != hipSuccess) return err.empty() ? hfail(ctx, e, "knn") : fail(ctx, OCT_ERR_ARG, err);
    if (keep) ctx->knn_nq = nq, ctx->knn_k = k;
    if (stats) {
        HCHK(hipStreamSynchronize(ctx->stream), "knn");
        if (int de = check_device_error(ctx)) return de;
        return read_stats(ctx, stats, stage_ms(ctx, ST_KNN));
    }
    return OCT_OK;
}

int oct_knn_fetch(oct_ctx *ctx, size_t q0, size_t nq, int32_t *idx_out, float *d2_out) {
    if (!ctx) return OCT_ERR_ARG;
    if (q0 > ctx->knn_nq || nq > ctx->knn_nq - q0) return fail(ctx, OCT_ERR_ARG, "knn fetch beyond the kept results");
    HCHK(hipSetDevice(ctx->device), "set device");
    const size_t k = ctx->knn_k, off = q0 * k, cnt = nq * k;
    if (idx_out && cnt) HCHK(hipMemcpyAsync(idx_out, ctx->kidx.p + off, cnt * 4, hipMemcpyDeviceToHost, ctx->stream), "copy idx");
    if (d2_out && cnt) HCHK(hipMemcpyAsync(d2_out, ctx->kd2f.p + off, cnt * 4, hipMemcpyDeviceToHost, ctx->stream), "copy d2");
    HCHK(hipStreamSynchronize(ctx->stream), "knn fetch");
    return check_device_error(ctx);
}

int oct_copy_bandwidth(oct_ctx *ctx, size_t bytes, double *ms_out) {
    if (!ctx || !ms_out) return OCT_ERR_ARG;
    if (bytes < 16 || bytes % 16) return fail(ctx, OCT_ERR_ARG, "copy size must be a positive multiple of 16 bytes");
    HCHK(hipSetDevice(ctx->device), "set device");
    HCHK(ctx->rec_in.reserve(bytes), "alloc copy source");
    HCHK(ctx->rec_out.reserve(bytes), "alloc copy target");
    HCHK(hipMemsetAsync(ctx->rec_in.p, 1, bytes, ctx->stream), "fill copy source");
    const size_t words = bytes / 16;
    if ((words + 255) / 256 > 0x7fffffffu) return fail(ctx, OCT_ERR_ARG, "copy size above 2^31 workgroups of 4 KiB");
    const unsigned nb = (unsigned)((words + 255) / 256);
    hipEvent_t e0, e1;
    HCHK(hipEventCreate(&e0), "event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, OCT_ERR_HIP, "event");
    }
    double best = 1e300;
    hipError_t err = hipSuccess;
    for (int rep = 0; rep < 6 && err == hipSuccess; rep++) {  // the first repetition warms up
        err = hipEventRecord(e0, ctx->stream);
        if (err == hipSuccess) {
            k_copy16<<<nb, 256, 0, ctx->stream>>>((const copy_word_t *)ctx->rec_in.p, (copy_word_t *)ctx->rec_out.p, words);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && rep) best = std::min(best, (double)ms);
    }
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    if (err != hipSuccess) return hfail(ctx, err, "copy");
    *ms_out = best;
    return OCT_OK;
}

int oct_knn_checksum(oct_ctx *ctx, const int32_t *idx_dev, const float *d2_dev, size_t nq, uint32_t k, uint64_t out[3]) {
    if (!ctx || !out) return OCT_ERR_ARG;
    if (!idx_dev != !d2_dev) return fail(ctx, OCT_ERR_ARG, "give both result buffers or neither");
    if (!idx_dev) {
        if (nq > ctx->knn_nq || (nq && k != ctx->knn_k)) return fail(ctx, OCT_ERR_ARG, "checksum beyond the kept results");
        idx_dev = ctx->kidx.p, d2_dev = ctx->kd2f.p;
    }
    HCHK(hipSetDevice(ctx->device), "set device");
    HCHK(ctx->stats.reserve(8), "alloc stats");
    HCHK(hipMemsetAsync(ctx->stats.p, 0, 3 * sizeof(unsigned long long), ctx->stream), "clear checksum");
    if (nq && k) {
        const unsigned nb = (unsigned)std::min<size_t>((nq + 255) / 256, 4096);
        k_knn_checksum<<<nb, 256, 0, ctx->stream>>>(idx_dev, d2_dev, nq, k, (uint32_t)ctx->n, ctx->stats.p);
        HCHK(hipGetLastError(), "knn checksum");
    }
    unsigned long long h[3];
    HCHK(hipMemcpyAsync(h, ctx->stats.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "copy checksum");
    HCHK(hipStreamSynchronize(ctx->stream), "knn checksum");
    for (int i = 0; i < 3; i++) out[i] = h[i];
    return check_device_error(ctx);
}

int oct_autotune_knn(oct_ctx *ctx, int precision, uint32_t k, size_t sample_nq, int *best, double ms_out[3]) {
    if (!ctx) return OCT_ERR_ARG;
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return fail(ctx, OCT_ERR_ARG, "unknown precision");
    if (!ctx->has_tree) return fail(ctx, OCT_ERR_STATE, "knn autotune before oct_build");
    if (k == 0 || k > 1024) return fail(ctx, OCT_ERR_ARG, "knn: k must be 1 .. 1024");
    if (ctx->n == 0) return fail(ctx, OCT_ERR_STATE, "autotune on an empty cloud");
    HCHK(hipSetDevice(ctx->device), "set device");
    const size_t nq = std::min(sample_nq ? sample_nq : (size_t)1 << 18, ctx->n);
    // 16 chunks of consecutive SFC positions spread over the cloud
    std::vector<uint32_t> idx(nq);
    const size_t chunks = 16, per = (nq + chunks - 1) / chunks;
    size_t n = 0;
    for (size_t c = 0; c < chunks && n < nq; c++) {
        const size_t start = (ctx->n / chunks) * c;
        for (size_t i = 0; i < per && n < nq; i++) idx[n++] = (uint32_t)std::min(start + i, ctx->n - 1);
    }
    HCHK(ctx->qidx.reserve(nq + 1), "alloc queries");
    HCHK(ctx->kidx.reserve(nq * k + 1), "alloc knn");
    HCHK(ctx->kd2f.reserve(nq * k + 1), "alloc knn");
    ctx->knn_nq = 0, ctx->knn_k = 0;
    HCHK(hipMemcpyAsync(ctx->qidx.p, idx.data(), nq * 4, hipMemcpyHostToDevice, ctx->stream), "copy sample");
    int opts[3];
    const int no = knn_lane_options(k, opts);
    double bestms = 1e300;
    int bests = opts[0];
    for (int o = 0; o < 3; o++) {
        if (ms_out) ms_out[o] = 1e300;
        if (o >= no) continue;
        KnnLaunch L;
        L.precision = precision;
        L.qidx = ctx->qidx.p;
        L.nq = nq, L.k = k;
        L.idx = ctx->kidx.p, L.d2f = ctx->kd2f.p;
        L.lanes = opts[o];
        double t = 1e300;
        for (int rep = 0; rep < 3; rep++) {  // first repetition warms caches and code objects
            std::string err;
            hipError_t e = launch_knn(ctx, L, err);
            if (e != hipSuccess) return err.empty() ? hfail(ctx, e, "knn autotune") : fail(ctx, OCT_ERR_ARG, err);
            HCHK(hipStreamSynchronize(ctx->stream), "knn autotune");
            if (rep) t = std::min(t, stage_ms(ctx, ST_KNN));
        }
        if (ms_out) ms_out[o] = t;
        if (t < bestms) bestms = t, bests = opts[o];
    }
    ctx->knn_lanes_tuned[precision & 1][knn_class(k)] = bests;
    if (best) *best = bests;
    return OCT_OK;
}

int oct_tuned_knn(oct_ctx *ctx, int precision, uint32_t k, int *lanes) {
    if (!ctx) return OCT_ERR_ARG;
    if (precision != OCT_PREC_FP64 && precision != OCT_PREC_FP32_RECHECK) return fail(ctx, OCT_ERR_ARG, "unknown precision");
    if (k == 0 || k > 1024) return fail(ctx, OCT_ERR_ARG, "knn: k must be 1 .. 1024");
    int opts[3];
    knn_lane_options(k, opts);
    const int l = knn_lanes(ctx, precision, k);
    if (lanes) *lanes = l > 0 ? l : opts[0];
    return OCT_OK;
}

int oct_stage_times(const oct_ctx *ctx, double ms[8]) {
    if (!ctx || !ms) return OCT_ERR_ARG;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return OCT_ERR_HIP;
    for (int s = 0; s < 8; s++) ms[s] = s < ST_N ? stage_ms(ctx, (Stage)s) : 0.0;
    return OCT_OK;
}

}  // extern "C"
