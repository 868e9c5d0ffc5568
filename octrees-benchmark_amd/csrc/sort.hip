// H3: stable LSD radix sort of (63-bit key, u32 index) pairs on gfx950, and the sorted-cloud gather.
//
// The reference sorts with 8 passes x 8 bits and per-thread histograms over contiguous chunks
// (point_encoder.hpp:293-376); its result is exactly "stable sort by key", which is what this sort
// produces. Design:
//  - digit histograms of every pass come from the encode kernel (no extra read of the keys);
//  - one kernel per pass ("onesweep"): a 256-thread workgroup takes a 6144-key tile (dynamic tile
//    ids, so tiles start in order), ranks keys stably with wave64 ballots on the 8 digit bits,
//    publishes per-bin tile counts and resolves its global bin offsets by decoupled look-back over
//    the preceding tiles, stages the tile in LDS in bin order and writes it out in runs;
//  - passes whose digit is the same for every key are skipped (the permutation is unchanged);
//  - every look-back wait is bounded: on timeout the kernel raises an error flag and still exits.
// Traffic per executed pass: 12 B/pt read + 12 B/pt written (+ 1 KB of look-back words per tile).
#include "device.hpp"

namespace oct {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 24;  // 6144-key tiles: 2 blocks per CU (80 KB LDS each); measured best of 8..32 keys per thread
static_assert(kSortBlock * kSortItems == kSortTile, "sort tile size");
constexpr uint32_t kFlagAgg = 1u << 30, kFlagInc = 2u << 30, kValMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 24;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    uint32_t l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(v, off);
        if (l >= (uint32_t)off) v += o;
    }
    return v;
}

// SEG = false: one key range [0, n), tile t covers [t * kSortTile, ...). SEG = true (batches): tiles
// come from a table, each inside one cloud's range; the look-back stops at the cloud's first tile and
// bin offsets are the cloud's own (bin_base + cloud * kSegBaseStride), so every cloud is sorted on
// its own, stably, in one launch.
template <bool SEG>
__global__ __launch_bounds__(kSortBlock) void k_onesweep(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, size_t n,
                                                         uint32_t shift, const uint32_t *__restrict__ bin_base,
                                                         uint32_t *__restrict__ lookback, uint32_t *__restrict__ tile_counter,
                                                         uint32_t *__restrict__ err, const SortTile *__restrict__ tiles) {
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_wave[kSortBlock / 64][256];
    __shared__ uint32_t s_tile_start[256];
    __shared__ uint32_t s_global[256];
    __shared__ uint32_t s_wsum[kSortBlock / 64];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (uint32_t i = tid; i < (kSortBlock / 64) * 256; i += kSortBlock) (&s_wave[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    size_t tile_base, limit;
    uint32_t first_tile = 0, cloud = 0;
    if constexpr (SEG) {
        const SortTile T = tiles[tile];
        tile_base = T.start, limit = (size_t)T.start + T.len, first_tile = T.first, cloud = T.cloud;
    } else {
        tile_base = (size_t)tile * kSortTile, limit = n;
    }
    const size_t wave_base = tile_base + (size_t)w * 64 * kSortItems;

    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t p = wave_base + (size_t)r * 64 + lane;
        bool valid = p < limit;
        k[r] = valid ? kin[p] : 0ull;
        v[r] = valid ? (vin ? vin[p] : (uint32_t)p) : 0u;
    }
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t p = wave_base + (size_t)r * 64 + lane;
        bool valid = p < limit;
        uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = ballot64(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t prev = s_wave[w][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) s_wave[w][d] = prev + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[r] = prev + below;
    }
    __syncthreads();

    // per-bin tile count and per-wave exclusive offsets (thread b < 256 owns bin b; with more than
    // 256 threads the other waves wait at the barriers)
    const uint32_t b = tid;
    const bool owner = kSortBlock == 256 || tid < 256;
    uint32_t tile_count = 0;
    uint32_t *lb = lookback + (size_t)tile * 256;
    if (owner) {
#pragma unroll
        for (int ww = 0; ww < kSortBlock / 64; ww++) {
            uint32_t c = s_wave[ww][b];
            s_wave[ww][b] = tile_count;
            tile_count += c;
        }
        if (tile == first_tile)
            store_flag(&lb[b], kFlagInc | tile_count);
        else
            store_flag(&lb[b], kFlagAgg | tile_count);
    }

    // local exclusive scan of the tile counts over bins
    uint32_t incl = 0;
    if (owner) {
        incl = wave_incl_scan(tile_count);
        if (lane == 63) s_wsum[w] = incl;
    }
    __syncthreads();
    if (owner) {
        uint32_t wpre = 0;
        for (uint32_t ww = 0; ww < w; ww++) wpre += s_wsum[ww];
        s_tile_start[b] = wpre + incl - tile_count;

        // decoupled look-back over preceding tiles for bin b
        uint32_t excl = 0;
        if (tile > first_tile) {
            int64_t t = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (t >= (int64_t)first_tile) {
                uint32_t x = load_flag(&lookback[(size_t)t * 256 + b]);
                if ((x >> 30) == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += x & kValMask;
                if ((x >> 30) == 2) break;
                t--;
            }
            store_flag(&lb[b], kFlagInc | ((excl + tile_count) & kValMask));
        }
        s_global[b] = bin_base[(size_t)cloud * kSegBaseStride + b] + excl;
    }  // owner
    __syncthreads();

    // stage the tile in LDS in (bin, original order) order
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t p = wave_base + (size_t)r * 64 + lane;
        if (p < limit) {
            uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
            uint32_t pos = s_tile_start[d] + s_wave[w][d] + rank[r];
            s_keys[pos] = k[r];
            s_vals[pos] = v[r];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (uint32_t)((limit - tile_base) < (size_t)kSortTile ? (limit - tile_base) : (size_t)kSortTile);
    for (uint32_t j = tid; j < tile_n; j += kSortBlock) {
        uint64_t key = s_keys[j];
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint32_t dst = s_global[d] + (j - s_tile_start[d]);
        kout[dst] = key;
        vout[dst] = s_vals[j];
    }
}

__global__ void k_iota_copy(const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                            size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        kout[i] = kin[i];
        vout[i] = (uint32_t)i;
    }
}

hipError_t launch_sort(oct_ctx *ctx, std::string &err) {
    const size_t n = ctx->n;
    if (n >= (size_t)kValMask) {
        err = "sort: n must be below 2^30";
        return hipErrorInvalidValue;
    }
    hipError_t e;
    if ((e = ctx->keys[1].reserve(n)) != hipSuccess) return e;
    if ((e = ctx->vals[0].reserve(n)) != hipSuccess) return e;
    if ((e = ctx->vals[1].reserve(n)) != hipSuccess) return e;
    // histograms of all passes (from the encode kernel)
    uint32_t h[8 * 256];
    if ((e = hipMemcpyAsync(h, ctx->hist.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    int passes[8], np = 0;
    uint32_t base[8 * 256];
    for (int p = 0; p < 8; p++) {
        uint32_t run = 0;
        bool trivial = false;
        for (int d = 0; d < 256; d++) {
            base[p * 256 + d] = run;
            run += h[p * 256 + d];
            if (h[p * 256 + d] == n) trivial = true;
        }
        if (!trivial && n > 1) passes[np++] = p;
    }
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    if ((e = ctx->lookback.reserve((size_t)(np ? np : 1) * tiles * 256 + 1)) != hipSuccess) return e;
    // misc layout: [0..16) tile counters, [16] error flag, [32..32+2048) bin bases
    if ((e = ctx->misc.reserve(12 * 1024 + 64)) != hipSuccess) return e;
    uint32_t *counters = ctx->misc.p, *errflag = ctx->misc.p + 16, *bases = ctx->misc.p + 32;
    stage_begin(ctx, ST_SORT);
    if ((e = hipMemcpyAsync(bases, base, sizeof(base), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(counters, 0, 32 * sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
    if (np && (e = hipMemsetAsync(ctx->lookback.p, 0, (size_t)np * tiles * 256 * sizeof(uint32_t), ctx->stream)) != hipSuccess)
        return e;
    int src = 0;
    for (int i = 0; i < np; i++) {
        int p = passes[i];
        const uint32_t *vin = (i == 0) ? nullptr : ctx->vals[src].p;
        k_onesweep<false><<<(unsigned)tiles, kSortBlock, 0, ctx->stream>>>(
            ctx->keys[src].p, vin, ctx->keys[src ^ 1].p, ctx->vals[src ^ 1].p, n, 8u * p, bases + p * 256,
            ctx->lookback.p + (size_t)i * tiles * 256, counters + i, errflag, nullptr);
        src ^= 1;
    }
    if (np == 0) {
        int nb = (int)((n + 255) / 256);
        nb = nb < 1 ? 1 : (nb > 4096 ? 4096 : nb);
        k_iota_copy<<<nb, 256, 0, ctx->stream>>>(ctx->keys[0].p, ctx->keys[1].p, ctx->vals[1].p, n);
        src = 1;
    }
    stage_end(ctx, ST_SORT);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    ctx->sorted_keys = ctx->keys[src].p;
    ctx->perm = ctx->vals[src].p;
    uint32_t flag = 0;
    if ((e = hipMemcpyAsync(&flag, errflag, sizeof(flag), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    if (flag) {
        err = "sort: look-back wait exceeded its bound";
        return hipErrorUnknown;
    }
    return hipSuccess;
}

hipError_t launch_sort_segmented(hipStream_t s, uint64_t *keys0, uint64_t *keys1, uint32_t *vals0, uint32_t *vals1,
                                 size_t n, const SortTile *tiles, size_t ntiles, const uint32_t *bases,
                                 uint32_t *lookback, uint32_t *counters, uint32_t *err) {
    if (n == 0 || ntiles == 0) return hipSuccess;
    uint64_t *kb[2] = {keys0, keys1};
    uint32_t *vb[2] = {vals0, vals1};
    for (int p = 0; p < kRadixPassCount; p++) {
        const int src = p & 1;
        k_onesweep<true><<<(unsigned)ntiles, kSortBlock, 0, s>>>(kb[src], p == 0 ? nullptr : vb[src], kb[src ^ 1], vb[src ^ 1], n,
                                                                8u * p, bases + p * 256, lookback + (size_t)p * ntiles * 256,
                                                                counters + p, err, tiles);
    }
    return hipGetLastError();
}

hipError_t launch_sort_passes(hipStream_t s, uint64_t *keys0, uint64_t *keys1, uint32_t *vals0, uint32_t *vals1, size_t n,
                              int npasses, const uint32_t *bases, uint32_t *lookback, uint32_t *counters, uint32_t *err) {
    if (n == 0) return hipSuccess;
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    uint64_t *kb[2] = {keys0, keys1};
    uint32_t *vb[2] = {vals0, vals1};
    for (int p = 0; p < npasses; p++) {
        const int src = p & 1;
        k_onesweep<false><<<(unsigned)tiles, kSortBlock, 0, s>>>(kb[src], p == 0 ? nullptr : vb[src], kb[src ^ 1], vb[src ^ 1], n,
                                                                 8u * p, bases + p * 256, lookback + (size_t)p * tiles * 256,
                                                                 counters + p, err, nullptr);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ gather of the sorted cloud
__global__ void k_gather(const double *__restrict__ xyz, const uint32_t *__restrict__ perm, double4 *__restrict__ out,
                         size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        size_t s = perm[i];
        out[i] = make_double4(xyz[3 * s], xyz[3 * s + 1], xyz[3 * s + 2], 0.0);
    }
}

// Records of any size moved by the permutation: out[i] = in[perm[i]], 16-byte words when the record
// size and both buffers allow it, else 4-byte words, else bytes.
template <class W>
__global__ void k_gather_records(const W *__restrict__ in, const uint32_t *__restrict__ perm, W *__restrict__ out, size_t n,
                                 uint32_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const W *src = in + (size_t)perm[i] * words;
        W *dst = out + i * words;
        for (uint32_t w = 0; w < words; w++) dst[w] = src[w];
    }
}

hipError_t launch_gather_records(hipStream_t s, const void *in, const uint32_t *perm, void *out, size_t n, size_t rb) {
    if (!n) return hipSuccess;
    int nb = (int)((n + 255) / 256);
    nb = nb > 8192 ? 8192 : nb;
    const uintptr_t al = (uintptr_t)in | (uintptr_t)out;
    if (rb % 16 == 0 && al % 16 == 0)
        k_gather_records<uint4><<<nb, 256, 0, s>>>((const uint4 *)in, perm, (uint4 *)out, n, (uint32_t)(rb / 16));
    else if (rb % 4 == 0 && al % 4 == 0)
        k_gather_records<uint32_t><<<nb, 256, 0, s>>>((const uint32_t *)in, perm, (uint32_t *)out, n, (uint32_t)(rb / 4));
    else
        k_gather_records<uint8_t><<<nb, 256, 0, s>>>((const uint8_t *)in, perm, (uint8_t *)out, n, (uint32_t)rb);
    return hipGetLastError();
}

// one point per thread over the whole grid: 5 % faster than a capped grid-stride grid at 50 M points
// (more rows in flight); n < 2^32 keeps the block count below 2^24
hipError_t launch_gather_raw(hipStream_t s, const double *xyz, const uint32_t *perm, double4 *out, size_t n) {
    if (!n) return hipSuccess;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_gather<<<nb, 256, 0, s>>>(xyz, perm, out, n);
    return hipGetLastError();
}

hipError_t launch_gather(oct_ctx *ctx) {
    hipError_t e;
    if ((e = ctx->p64.reserve(ctx->n)) != hipSuccess) return e;
    stage_begin(ctx, ST_GATHER);
    if (ctx->n) {
        const unsigned nb = (unsigned)((ctx->n + 255) / 256);  // one point per thread (launch_gather_raw)
        k_gather<<<nb, 256, 0, ctx->stream>>>(ctx->xyz.p, ctx->perm, ctx->p64.p, ctx->n);
    }
    stage_end(ctx, ST_GATHER);
    return hipGetLastError();
}

// ------------------------------------------------------------------ exclusive scans (3-phase)
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

template <class T>
__device__ inline T block_excl_scan(T v, T *s_w, T &total) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    T incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T o = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += o;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    T pre = 0, tot = 0;
    for (uint32_t ww = 0; ww < kScanBlock / 64; ww++) {
        if (ww < w) pre += s_w[ww];
        tot += s_w[ww];
    }
    __syncthreads();
    total = tot;
    return pre + incl - v;
}

template <class TI>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t *__restrict__ in, size_t n, TI *__restrict__ part) {
    __shared__ TI s_w[kScanBlock / 64];
    size_t base = (size_t)blockIdx.x * kScanTile;
    TI s = 0;
    for (int i = 0; i < kScanItems; i++) {
        size_t p = base + (size_t)i * kScanBlock + threadIdx.x;
        if (p < n) s += in[p];
    }
    TI tot;
    (void)block_excl_scan<TI>(s, s_w, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <class TI>
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(TI *__restrict__ part, size_t nb, TI *__restrict__ total) {
    __shared__ TI s_w[kScanBlock / 64];
    TI carry = 0;
    for (size_t base = 0; base < nb; base += kScanBlock) {
        size_t p = base + threadIdx.x;
        TI v = p < nb ? part[p] : 0;
        TI tot;
        TI ex = block_excl_scan<TI>(v, s_w, tot);
        if (p < nb) part[p] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <class TI, class TO>
__global__ __launch_bounds__(kScanBlock) void k_scan_down(const uint32_t *__restrict__ in, size_t n, const TI *__restrict__ part,
                                                         TO *__restrict__ out) {
    __shared__ TI s_w[kScanBlock / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    TI v[kScanItems], s = 0;
    for (int i = 0; i < kScanItems; i++) {
        size_t p = base + i;
        v[i] = p < n ? in[p] : 0;
        s += v[i];
    }
    TI tot;
    TI run = part[blockIdx.x] + block_excl_scan<TI>(s, s_w, tot);
    for (int i = 0; i < kScanItems; i++) {
        size_t p = base + i;
        if (p < n) out[p] = (TO)run;
        run += v[i];
    }
}

template <class TI, class TO>
static hipError_t scan_impl(oct_ctx *ctx, const uint32_t *in, TO *out, size_t n, TI *total_dev) {
    size_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) nb = 1;
    hipError_t e = ctx->scan.reserve(2 * nb * sizeof(TI) / sizeof(uint32_t) + 8);
    if (e != hipSuccess) return e;
    TI *part = reinterpret_cast<TI *>(ctx->scan.p);
    k_scan_reduce<TI><<<(unsigned)nb, kScanBlock, 0, ctx->stream>>>(in, n, part);
    k_scan_partials<TI><<<1, kScanBlock, 0, ctx->stream>>>(part, nb, total_dev);
    k_scan_down<TI, TO><<<(unsigned)nb, kScanBlock, 0, ctx->stream>>>(in, n, part, out);
    return hipGetLastError();
}

hipError_t exclusive_scan_u32_to_u64(oct_ctx *ctx, const uint32_t *in, uint64_t *out, size_t n) {
    // out has n + 1 entries; out[n] = total
    return scan_impl<uint64_t, uint64_t>(ctx, in, out, n, out + n);
}

hipError_t exclusive_scan_u32(oct_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev) {
    return scan_impl<uint32_t, uint32_t>(ctx, in, out, n, total_dev);
}

}  // namespace oct
