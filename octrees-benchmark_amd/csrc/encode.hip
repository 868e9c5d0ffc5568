// H1 + H2: bounding box reduction and SFC encoding (gfx950).
//
// k_bbox:   FP64 min/max per axis (box.hpp:94-137); 24 B/pt streamed once.
// k_encode: per point FP64 quantisation (point_encoder.hpp:37-49) + Morton / Hilbert key, and the
//           digit histograms of all radix passes, and of bits 55..62 and 47..54 for the partition
//           sort, in the same read (so the sort needs no histogram pass). 24 B/pt read + 8 B/pt written.
#include <cfloat>

#include "device.hpp"

namespace oct {

constexpr int kBboxBlock = 256;
constexpr int kBboxBlocks = 1024;

__global__ __launch_bounds__(kBboxBlock) void k_bbox(const double *__restrict__ xyz, size_t n,
                                                     double *__restrict__ partial) {
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; a++) {
            double v = xyz[3 * i + a];
            mn[a] = (v < mn[a]) ? v : mn[a];  // std::min semantics (NaN never replaces the running value)
            mx[a] = (mx[a] < v) ? v : mx[a];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; a++) {
            double o = __shfl_xor(mn[a], off);
            mn[a] = (o < mn[a]) ? o : mn[a];
            o = __shfl_xor(mx[a], off);
            mx[a] = (mx[a] < o) ? o : mx[a];
        }
    }
    __shared__ double s[kBboxBlock / 64][6];
    int w = threadIdx.x / 64;
    if (lane_id() == 0)
        for (int a = 0; a < 3; a++) s[w][a] = mn[a], s[w][3 + a] = mx[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < kBboxBlock / 64; j++)
            for (int a = 0; a < 3; a++) {
                s[0][a] = (s[j][a] < s[0][a]) ? s[j][a] : s[0][a];
                s[0][3 + a] = (s[0][3 + a] < s[j][3 + a]) ? s[j][3 + a] : s[0][3 + a];
            }
        for (int k = 0; k < 6; k++) partial[(size_t)blockIdx.x * 6 + k] = s[0][k];
    }
}

// out: min and max per axis; with box, also [min, max, centre, radii] per axis there, the centre and
// radii in the reference's mbb expressions (util.hpp:139-143), so they equal the host's values.
__global__ __launch_bounds__(kBboxBlock) void k_bbox_final(const double *__restrict__ partial, int nb,
                                                           double *__restrict__ out, double *__restrict__ box) {
    __shared__ double s[kBboxBlock / 64][6];
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int a = 0; a < 3; a++) {
            double lo = partial[(size_t)b * 6 + a], hi = partial[(size_t)b * 6 + 3 + a];
            v[a] = (lo < v[a]) ? lo : v[a];
            v[3 + a] = (v[3 + a] < hi) ? hi : v[3 + a];
        }
    // the partials hold no NaN (a NaN never replaces a running value), so the order of the
    // combination does not change the result: lanes by exchange, then the block's waves
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++) {
            const double lo = __shfl_xor(v[a], off), hi = __shfl_xor(v[3 + a], off);
            v[a] = (lo < v[a]) ? lo : v[a];
            v[3 + a] = (v[3 + a] < hi) ? hi : v[3 + a];
        }
    const int w = threadIdx.x / 64;
    if (lane_id() == 0)
        for (int k = 0; k < 6; k++) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < kBboxBlock / 64; j++)
            for (int a = 0; a < 3; a++) {
                v[a] = (s[j][a] < v[a]) ? s[j][a] : v[a];
                v[3 + a] = (v[3 + a] < s[j][3 + a]) ? s[j][3 + a] : v[3 + a];
            }
        for (int k = 0; k < 6; k++) out[k] = v[k];
        if (box)
            for (int a = 0; a < 3; a++) {
                box[a] = v[a], box[3 + a] = v[3 + a];
                box[6 + a] = v[a] + (v[3 + a] - v[a]) / 2;
                box[9 + a] = (v[3 + a] - v[a]) / 2;
            }
    }
}

static hipError_t enqueue_bbox(oct_ctx *ctx, double **out, double *box) {
    hipError_t e = ctx->misc.reserve(kBboxBlocks * 12 + 64);
    if (e != hipSuccess) return e;
    double *partial = reinterpret_cast<double *>(ctx->misc.p);
    *out = partial + (size_t)kBboxBlocks * 6;
    int nb = (int)((ctx->n + kBboxBlock - 1) / kBboxBlock);
    nb = nb < 1 ? 1 : (nb > kBboxBlocks ? kBboxBlocks : nb);
    stage_begin(ctx, ST_BBOX);
    k_bbox<<<nb, kBboxBlock, 0, ctx->stream>>>(ctx->xyz.p, ctx->n, partial);
    k_bbox_final<<<1, kBboxBlock, 0, ctx->stream>>>(partial, nb, *out, box);
    stage_end(ctx, ST_BBOX);
    return hipGetLastError();
}

hipError_t launch_bbox_dev(oct_ctx *ctx) {
    hipError_t e = ctx->box.reserve(12);
    if (e != hipSuccess) return e;
    double *out = nullptr;
    return enqueue_bbox(ctx, &out, ctx->box.p);
}

hipError_t launch_bbox(oct_ctx *ctx, double mn[3], double mx[3]) {
    double *out = nullptr;
    hipError_t e = enqueue_bbox(ctx, &out, nullptr);
    if (e != hipSuccess) return e;
    double h[6];
    e = hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    for (int a = 0; a < 3; a++) mn[a] = h[a], mx[a] = h[3 + a];
    return hipGetLastError();
}

// ------------------------------------------------------------------ encode + digit histograms
constexpr int kEncBlock = 256;
constexpr int kRadixPasses = 8;
static_assert(kRadixPasses == kRadixPassCount, "one histogram per LSD pass");

struct EncodeBox {
    double c[3], r[3];
};

// LSDH: the 8 LSD digit histograms and the partition passes' (bits 55..62, 47..54); else only the
// latter (the partition sort needs no other, and two LDS atomics per point are cheaper than ten).
// The box: box_arg, or with boxd (min, max, centre, radii per axis, as k_bbox_final leaves them) the
// centre and radii read from there.
template <int ENC, bool LSDH>
__global__ __launch_bounds__(kEncBlock) void k_encode(const double *__restrict__ xyz, size_t n, EncodeBox box_arg,
                                                      const double *__restrict__ boxd, uint64_t *__restrict__ keys,
                                                      uint32_t *__restrict__ hist) {
    EncodeBox box = box_arg;
    if (boxd)
        for (int a = 0; a < 3; a++) box.c[a] = boxd[6 + a], box.r[a] = boxd[9 + a];
    __shared__ uint32_t sh[kEncHists * 256];
    __shared__ uint16_t stab[48 * 8];
    __shared__ uint16_t stab2[kHilbertTab2];
    for (int i = threadIdx.x; i < kEncHists * 256; i += blockDim.x) sh[i] = 0;
    if (ENC == OCT_ENC_HILBERT3D)
        for (int i = threadIdx.x; i < 48 * 8; i += blockDim.x) stab[i] = kHilbert.t[i];
    __syncthreads();
    if (ENC == OCT_ENC_HILBERT3D) {
        hilbert_tab2_fill(stab2, stab, threadIdx.x, blockDim.x);
        __syncthreads();
    }
    auto point = [&](size_t i, double px, double py, double pz) {
        const uint32_t x = quantize_axis(px, box.c[0], box.r[0]);
        const uint32_t y = quantize_axis(py, box.c[1], box.r[1]);
        const uint32_t z = quantize_axis(pz, box.c[2], box.r[2]);
        const uint64_t k = (ENC == OCT_ENC_HILBERT3D) ? hilbert_encode_tab2(stab, stab2, x, y, z) : morton_encode(x, y, z);
        keys[i] = k;
        if (LSDH) {
#pragma unroll
            for (int p = 0; p < kRadixPasses; p++) atomicAdd(&sh[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
        }
        atomicAdd(&sh[kRadixPasses * 256 + ((k >> kMsdShift1) & 255u)], 1u);
        atomicAdd(&sh[kRadixPasses * 256 + 256 + ((k >> kMsdShift2) & 255u)], 1u);
    };
    // two points per thread and iteration: both points' loads are in flight together and their
    // table walks interleave
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
        const size_t i2 = i + stride;
        const double x0 = xyz[3 * i], y0 = xyz[3 * i + 1], z0 = xyz[3 * i + 2];
        const double x1 = xyz[3 * i2], y1 = xyz[3 * i2 + 1], z1 = xyz[3 * i2 + 2];
        point(i, x0, y0, z0);
        point(i2, x1, y1, z1);
    }
    if (i < n) point(i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    __syncthreads();
    for (int i = threadIdx.x; i < kEncHists * 256; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// Batched key algebra for single coordinates (PointEncoder::encode / decode,
// morton_encoder_3d.hpp:35-46, hilbert_encoder_3d.hpp:29-66, 256-289).
template <int ENC>
__global__ __launch_bounds__(256) void k_key_codec(uint32_t *__restrict__ xyz, uint64_t *__restrict__ keys, size_t n,
                                                  bool decode) {
    __shared__ uint16_t s_tab[48 * 8];
    if (ENC == OCT_ENC_HILBERT3D && !decode)
        for (int t = threadIdx.x; t < 48 * 8; t += blockDim.x) s_tab[t] = kHilbert.t[t];
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (decode) {
            uint32_t x, y, z;
            if (ENC == OCT_ENC_HILBERT3D)
                hilbert_decode(keys[i], x, y, z);
            else
                morton_decode(keys[i], x, y, z);
            xyz[3 * i] = x, xyz[3 * i + 1] = y, xyz[3 * i + 2] = z;
        } else {
            const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            keys[i] = ENC == OCT_ENC_HILBERT3D ? hilbert_encode_tab(s_tab, x, y, z) : morton_encode(x, y, z);
        }
    }
}

hipError_t launch_key_codec(oct_ctx *ctx, int encoder, bool decode, uint32_t *xyz_dev, uint64_t *keys_dev, size_t n) {
    if (!n) return hipSuccess;
    unsigned nb = (unsigned)((n + 255) / 256);
    nb = nb > 2048 ? 2048 : nb;
    if (encoder == OCT_ENC_HILBERT3D)
        k_key_codec<OCT_ENC_HILBERT3D><<<nb, 256, 0, ctx->stream>>>(xyz_dev, keys_dev, n, decode);
    else
        k_key_codec<OCT_ENC_MORTON3D><<<nb, 256, 0, ctx->stream>>>(xyz_dev, keys_dev, n, decode);
    return hipGetLastError();
}

hipError_t launch_encode(oct_ctx *ctx, int encoder, const double center[3], const double radii[3]) {
    size_t n = ctx->n;
    hipError_t e;
    if ((e = ctx->keys[0].reserve(n)) != hipSuccess) return e;
    if ((e = ctx->hist.reserve(kEncHists * 256)) != hipSuccess) return e;
    EncodeBox box = {};
    const double *boxd = center ? nullptr : ctx->box.p;
    if (!center && !boxd) return hipErrorInvalidValue;  // no box on the host, none left on the device
    if (center)
        for (int a = 0; a < 3; a++) box.c[a] = center[a], box.r[a] = radii[a];
    stage_begin(ctx, ST_ENCODE);
    e = hipMemsetAsync(ctx->hist.p, 0, kEncHists * 256 * sizeof(uint32_t), ctx->stream);
    if (e != hipSuccess) return e;
    if (n) {
        int nb = (int)((n + kEncBlock * 16 - 1) / (kEncBlock * 16));
        nb = nb < 1 ? 1 : (nb > 2048 ? 2048 : nb);
        const bool lsdh = !sort_partition_planned(ctx);
        if (encoder == OCT_ENC_HILBERT3D) {
            if (lsdh)
                k_encode<OCT_ENC_HILBERT3D, true><<<nb, kEncBlock, 0, ctx->stream>>>(ctx->xyz.p, n, box, boxd, ctx->keys[0].p, ctx->hist.p);
            else
                k_encode<OCT_ENC_HILBERT3D, false><<<nb, kEncBlock, 0, ctx->stream>>>(ctx->xyz.p, n, box, boxd, ctx->keys[0].p, ctx->hist.p);
        } else {
            if (lsdh)
                k_encode<OCT_ENC_MORTON3D, true><<<nb, kEncBlock, 0, ctx->stream>>>(ctx->xyz.p, n, box, boxd, ctx->keys[0].p, ctx->hist.p);
            else
                k_encode<OCT_ENC_MORTON3D, false><<<nb, kEncBlock, 0, ctx->stream>>>(ctx->xyz.p, n, box, boxd, ctx->keys[0].p, ctx->hist.p);
        }
        ctx->hist_lsd = lsdh;
    }
    stage_end(ctx, ST_ENCODE);
    return hipGetLastError();
}

}  // namespace oct
