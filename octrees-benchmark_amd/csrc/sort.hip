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
#include <algorithm>

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

// SEG = false: one key range [0, n), tile t covers [t * kSortTile, ...). SEG = true (batches, and the
// second partition pass below): tiles come from a table, each inside one segment's range; the
// look-back stops at the segment's first tile and bin offsets are the segment's own
// (bin_base + segment * base_stride), so every segment is sorted on its own, stably, in one launch.
template <bool SEG>
__global__ __launch_bounds__(kSortBlock) void k_onesweep(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, size_t n,
                                                         uint32_t shift, const uint32_t *__restrict__ bin_base,
                                                         uint32_t *__restrict__ lookback, uint32_t *__restrict__ tile_counter,
                                                         uint32_t *__restrict__ err, const SortTile *__restrict__ tiles,
                                                         uint32_t base_stride, uint32_t *__restrict__ hist16) {
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
        // the lowest lane of each digit group adds the group to the wave's digit count and hands the
        // count before it to its group; a wave's LDS atomics run in program order, so rows rank in order
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t prev = 0;
        if (valid && below == 0) prev = atomicAdd(&s_wave[w][d], (uint32_t)__popcll(peers));
        const uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : 0u;
        rank[r] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)prev) + below;
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
        s_global[b] = bin_base[(size_t)cloud * base_stride + b] + excl;
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
        if (hist16) {
            // the keys' top 16 bits (47..62) per prefix: when this pass orders bits 55..62 after a pass
            // on 47..54, the staged tile is ordered by them, so each run of one prefix adds its length
            // (its first key subtracts its index, its last adds index + 1)
            const uint32_t pre = (uint32_t)(key >> kMsdShift2);
            if (j == 0 || (uint32_t)(s_keys[j - 1] >> kMsdShift2) != pre) atomicAdd(&hist16[pre], 0u - j);
            if (j + 1 == tile_n || (uint32_t)(s_keys[j + 1] >> kMsdShift2) != pre) atomicAdd(&hist16[pre], j + 1u);
        }
    }
}

__global__ void k_iota_copy(const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                            size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        kout[i] = kin[i];
        vout[i] = (uint32_t)i;
    }
}

// ------------------------------------------------------------------ partition sort
//
// For more keys than one workgroup sorts in LDS, all below 2^63 (both 3-D encoders' keys), the same
// stable sort runs as two device-wide passes and one workgroup pass instead of eight device-wide passes:
//  1. the onesweep pass above on bits 47..54, then on bits 55..62 (both histograms come from the encode
//     kernel): the keys are now ordered by their top 16 bits. The second pass stages each tile in that
//     order, so it also counts the keys of every 16-bit prefix (two atomics per run of one prefix);
//  2. k_msd_jobs turns those counts into workgroup jobs: one 16-bit segment larger than the
//     workgroup's LDS capacity, or a run of consecutive smaller segments of one top-8-bit bin that fits;
//  3. k_seg_sort: one workgroup per job sorts its range completely. A range that fits is sorted in LDS
//     by LSD passes over only the digits, and inside a digit only the bits, that vary in it, each
//     ranked stably with wave ballots. A larger segment is first split stably by its next 8 bits
//     (39..46, then 31..38, ...) in LDS-sized tiles with running bin offsets (one workgroup owns the
//     segment, so no look-back), until every piece fits. The workgroup then reads the points of its
//     sorted range through the permutation and writes the sorted cloud (the gather), so the random
//     reads overlap other workgroups' LDS work.
// Every step keeps equal keys in input order, so the permutation is the stable sort by key, as with
// the 8 LSD passes. Traffic per key: 20 B + 24 B (passes) + 24 B (LDS sort) + 56 B (gather), plus
// 32 B for keys in segments that are split, against 188 B + 60 B for eight passes and a gather.

struct SegJob {
    uint32_t start, len;
};
constexpr uint32_t kSegMaxLen = 1u << 15;  // longest 16-bit segment the workgroup sort takes
// pieces awaiting a split: a split of up to 2^15 keys pushes at most 15 pieces above 2048 keys and
// pops one, over at most 6 levels (bits 39..46 down to 0..7)
constexpr int kSegStack = 96;

template <int W, int IPT>
struct SegShared {
    static constexpr uint32_t kCap = (uint32_t)(W * 64 * IPT);
    uint64_t keys[kCap];
    uint32_t vals[kCap];
    uint32_t cnt[W][256];  // per wave: digit counts, then the wave's first position inside each digit
    uint32_t start[256];   // first position of each digit in the ranked range
    uint32_t tot[256];     // keys per digit in the ranked range
    uint32_t hist[256];    // split: keys per digit of the whole piece, then the start of each run to sort
    uint32_t off[256];     // split: next output position of each digit, then the length of each run
    uint32_t wsum[W];
    uint32_t mask_lo, mask_hi;  // OR of (key ^ first key) over a range
    uint32_t or_d, and_d;       // split: OR / AND of the digit values present
    uint32_t sp, nchunks, kind;
    SegJob job;
    uint4 stack[kSegStack];  // start, len, shift + 1 (0: no bits left, keys equal), buffer (0: a, 1: b)
};

// Greedy packing, by one wave, of 256 consecutive digit ranges (digit d = 4 * lane + slot: count
// c[slot], start st[slot]), in digit order: a range longer than cap is reported alone (on_big), runs of
// consecutive shorter ranges up to cap together (on_run); empty digits are skipped. Every lane runs the
// same scalar loop, so the callbacks are wave-uniform.
template <class OnBig, class OnRun>
__device__ inline void pack_digit_ranges(const uint32_t (&c)[4], const uint32_t (&st)[4], uint32_t cap, OnBig &&on_big,
                                         OnRun &&on_run) {
    uint64_t nz[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) nz[sl] = ballot64(c[sl] != 0u);
    uint32_t rs = 0, rl = 0;
    for (uint32_t ln = 0; ln < 64; ln++) {
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
            if (!((nz[sl] >> ln) & 1u)) continue;
            const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)c[sl], (int)ln);
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)st[sl], (int)ln);
            if (l > cap) {
                if (rl) on_run(rs, rl), rl = 0;
                on_big(a, l);
            } else {
                if (rl + l > cap) on_run(rs, rl), rl = 0;
                if (!rl) rs = a;
                rl += l;
            }
        }
    }
    if (rl) on_run(rs, rl);
}

// Stable rank of the n keys a workgroup holds (thread (w, lane), row r: position w*64*IPT + r*64 + lane)
// by the digit (key >> shift) & 255, with ballots over the digit bits in dbits only (the others are
// equal for all n keys). Afterwards S.cnt[w][d] + rank[r] is the key's position in digit order,
// S.start[d] the first position of digit d and S.tot[d] the number of keys with digit d.
template <int W, int IPT>
__device__ inline void seg_rank(SegShared<W, IPT> &S, const uint64_t (&k)[IPT], uint32_t n, uint32_t shift,
                                uint32_t dbits, uint32_t (&rank)[IPT]) {
    static_assert(W >= 4, "256 threads own the 256 digits");
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63, base = w * 64 * IPT;
    for (uint32_t i = lane; i < 256; i += 64) S.cnt[w][i] = 0u;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        if (base + (uint32_t)r * 64 >= n) break;
        const bool valid = base + (uint32_t)r * 64 + lane < n;
        const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint64_t peers = ballot64(valid);
        for (uint32_t bm = dbits; bm; bm &= bm - 1u) {
            const bool bit = (d >> __builtin_ctz(bm)) & 1u;
            const uint64_t m = ballot64(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t prev = 0;
        if (valid && below == 0) prev = atomicAdd(&S.cnt[w][d], (uint32_t)__popcll(peers));
        const uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : 0u;
        rank[r] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)prev) + below;
    }
    __syncthreads();
    uint32_t c[W], tot = 0, incl = 0;
    if (tid < 256) {
#pragma unroll
        for (int ww = 0; ww < W; ww++) c[ww] = S.cnt[ww][tid], tot += c[ww];
        S.tot[tid] = tot;
        incl = wave_incl_scan(tot);
        if (lane == 63) S.wsum[w] = incl;
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t pre = 0;
        for (uint32_t ww = 0; ww < w; ww++) pre += S.wsum[ww];
        uint32_t st = pre + incl - tot;
        S.start[tid] = st;
#pragma unroll
        for (int ww = 0; ww < W; ww++) S.cnt[ww][tid] = st, st += c[ww];
    }
    __syncthreads();
}

// Stable sort of [start, start + n) of (ks, vs), 1 <= n <= kCap, into the same range of (kd, vd), which
// may be the source: LSD passes over only the 8-bit digits, and inside a digit only the bits, that
// vary in the range. With xyz, also writes the sorted cloud's records of the range to p64. The caller
// synchronises the workgroup after it.
template <int W, int IPT>
__device__ void seg_local_sort(SegShared<W, IPT> &S, const uint64_t *ks, const uint32_t *vs, uint64_t *kd, uint32_t *vd,
                               uint32_t start, uint32_t n, const double *xyz, double4 *p64) {
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63, base = w * 64 * IPT;
    uint64_t k[IPT];
    uint32_t v[IPT], rank[IPT];
    if (tid == 0) S.mask_lo = S.mask_hi = 0u;
    const uint64_t k0 = ks[start];
    uint64_t m = 0;
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const uint32_t p = base + (uint32_t)r * 64 + lane;
        k[r] = p < n ? ks[start + p] : 0ull;
        v[r] = p < n ? vs[start + p] : 0u;
        if (p < n) m |= k[r] ^ k0;
    }
    uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
    lo |= lane_xor<1>(lo), hi |= lane_xor<1>(hi);
    lo |= lane_xor<2>(lo), hi |= lane_xor<2>(hi);
    lo |= lane_xor<4>(lo), hi |= lane_xor<4>(hi);
    lo |= lane_xor<8>(lo), hi |= lane_xor<8>(hi);
    lo |= lane_xor<16>(lo), hi |= lane_xor<16>(hi);
    lo |= lane_xor32(lo), hi |= lane_xor32(hi);
    __syncthreads();
    if (lane == 0) {
        if (lo) atomicOr(&S.mask_lo, lo);
        if (hi) atomicOr(&S.mask_hi, hi);
    }
    __syncthreads();
    const uint64_t mask = ((uint64_t)S.mask_hi << 32) | S.mask_lo;
    for (uint32_t sh = 0; sh < 64; sh += 8) {
        const uint32_t dbits = (uint32_t)(mask >> sh) & 255u;
        if (!dbits) continue;
        seg_rank(S, k, n, sh, dbits, rank);
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const uint32_t p = base + (uint32_t)r * 64 + lane;
            if (p < n) {
                const uint32_t pos = S.cnt[w][(uint32_t)(k[r] >> sh) & 255u] + rank[r];
                S.keys[pos] = k[r];
                S.vals[pos] = v[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const uint32_t p = base + (uint32_t)r * 64 + lane;
            if (p < n) k[r] = S.keys[p], v[r] = S.vals[p];
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const uint32_t p = base + (uint32_t)r * 64 + lane;
        if (p < n) kd[start + p] = k[r], vd[start + p] = v[r];
    }
    if (xyz) {  // the sorted cloud's records of the range, read through the permutation
        double x[IPT], y[IPT], z[IPT];
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            if (base + (uint32_t)r * 64 + lane < n) {
                const double *q = xyz + 3 * (size_t)v[r];
                x[r] = q[0], y[r] = q[1], z[r] = q[2];
            }
        }
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const uint32_t p = base + (uint32_t)r * 64 + lane;
            if (p < n) p64[start + p] = make_double4(x[r], y[r], z[r], 0.0);
        }
    }
}

// Complete sort of one 16-bit segment longer than kCap, [start, start + len) of buffer a, into the
// same range of buffer b (buffer a's range is scratch): stable splits by the next 8 bits until every
// piece fits, then LDS sorts of runs of consecutive pieces. A piece whose bits are used up holds
// equal keys, already in input order.
template <int W, int IPT>
__device__ void seg_split_sort(SegShared<W, IPT> &S, uint64_t *ka, uint32_t *va, uint64_t *kb, uint32_t *vb,
                               uint32_t start, uint32_t len, uint32_t *err, const double *xyz, double4 *p64) {
    constexpr uint32_t kCap = SegShared<W, IPT>::kCap;
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63, base = w * 64 * IPT;
    if (tid == 0) {
        S.stack[0] = make_uint4(start, len, kMsdShift3 + 1u, 0u);
        S.sp = 1u;
    }
    __syncthreads();
    for (;;) {
        const uint32_t sp = S.sp;
        if (sp == 0) break;
        const uint4 e = S.stack[sp - 1];
        __syncthreads();
        if (tid == 0) S.sp = sp - 1;
        uint64_t *ks = e.w ? kb : ka, *kt = e.w ? ka : kb;
        uint32_t *vs = e.w ? vb : va, *vt = e.w ? va : vb;
        if (e.z == 0) {
            for (uint32_t i = tid; i < e.y; i += W * 64) {
                const uint32_t val = e.w ? vb[e.x + i] : va[e.x + i];
                if (e.w == 0) kb[e.x + i] = ka[e.x + i], vb[e.x + i] = val;
                if (xyz) {
                    const double *q = xyz + 3 * (size_t)val;
                    p64[e.x + i] = make_double4(q[0], q[1], q[2], 0.0);
                }
            }
            __syncthreads();
            continue;
        }
        const uint32_t shift = e.z - 1u;
        const uint32_t next1 = shift >= 8u ? shift - 7u : (shift > 0u ? 1u : 0u);  // next shift + 1
        if (tid < 256) S.hist[tid] = 0u;
        if (tid == 0) S.or_d = 0u, S.and_d = 255u;
        __syncthreads();
        for (uint32_t t0 = 0; t0 < e.y; t0 += kCap) {
            uint32_t dg[IPT];
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const uint32_t p = t0 + base + (uint32_t)r * 64 + lane;
                dg[r] = p < e.y ? (uint32_t)(ks[e.x + p] >> shift) & 255u : 256u;
            }
#pragma unroll
            for (int r = 0; r < IPT; r++)
                if (dg[r] < 256u) atomicAdd(&S.hist[dg[r]], 1u);
        }
        __syncthreads();
        uint32_t c = 0, incl = 0;
        if (tid < 256) {
            c = S.hist[tid];
            if (c) atomicOr(&S.or_d, tid), atomicAnd(&S.and_d, tid);
            incl = wave_incl_scan(c);
            if (lane == 63) S.wsum[w] = incl;
        }
        __syncthreads();
        if (tid < 256) {
            uint32_t pre = 0;
            for (uint32_t ww = 0; ww < w; ww++) pre += S.wsum[ww];
            S.off[tid] = e.x + pre + incl - c;
        }
        __syncthreads();
        const uint32_t dbits = S.or_d & ~S.and_d;
        if (dbits == 0) {  // one digit value: nothing moves, go on with the next bits
            if (tid == 0) S.stack[sp - 1] = make_uint4(e.x, e.y, next1, e.w), S.sp = sp;
            __syncthreads();
            continue;
        }
        for (uint32_t t0 = 0; t0 < e.y; t0 += kCap) {
            const uint32_t tn = e.y - t0 < kCap ? e.y - t0 : kCap;
            uint64_t k[IPT];
            uint32_t v[IPT], rank[IPT];
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const uint32_t p = base + (uint32_t)r * 64 + lane;
                k[r] = p < tn ? ks[e.x + t0 + p] : 0ull;
                v[r] = p < tn ? vs[e.x + t0 + p] : 0u;
            }
            seg_rank(S, k, tn, shift, dbits, rank);
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const uint32_t p = base + (uint32_t)r * 64 + lane;
                if (p < tn) {
                    const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
                    const uint32_t pos = S.cnt[w][d] + rank[r];
                    S.keys[pos] = k[r];
                    S.vals[pos] = v[r];
                }
            }
            __syncthreads();
            for (uint32_t j = tid; j < tn; j += W * 64) {
                const uint64_t key = S.keys[j];
                const uint32_t d = (uint32_t)(key >> shift) & 255u;
                const uint32_t dst = S.off[d] + (j - S.start[d]);
                kt[dst] = key;
                vt[dst] = S.vals[j];
            }
            __syncthreads();
            if (tid < 256) S.off[tid] += S.tot[tid];
            __syncthreads();
        }
        // pieces: digit d holds [off[d] - hist[d], off[d]) of buffer t; big ones go on the stack, runs
        // of small ones become LDS sorts
        if (w == 0) {
            uint32_t c4[4], s4[4];
#pragma unroll
            for (int sl = 0; sl < 4; sl++) c4[sl] = S.hist[4 * lane + sl], s4[sl] = S.off[4 * lane + sl] - c4[sl];
            uint32_t nc = 0, top = sp - 1;
            pack_digit_ranges(
                c4, s4, kCap,
                [&](uint32_t a, uint32_t l) {
                    if (top < (uint32_t)kSegStack) {
                        if (lane == 0) S.stack[top] = make_uint4(a, l, next1, e.w ^ 1u);
                        top++;
                    } else if (lane == 0) {
                        atomicOr(err, 2u);
                    }
                },
                [&](uint32_t a, uint32_t l) {  // wave 0 has read hist / off already
                    if (lane == 0) S.hist[nc] = a, S.off[nc] = l;
                    nc++;
                });
            if (lane == 0) S.nchunks = nc, S.sp = top;
        }
        __syncthreads();
        const uint32_t nc = S.nchunks;
        for (uint32_t ci = 0; ci < nc; ci++) {  // the LDS sorts leave hist / off alone
            seg_local_sort(S, kt, vt, kb, vb, S.hist[ci], S.off[ci], xyz, p64);
            __syncthreads();
        }
    }
}

// Persistent workgroups over the jobs of k_msd_jobs (the big ones first): the second partition
// pass's output (ka, va) in, the sorted keys and permutation (kb, vb) out, and with xyz (n x 3 input
// points) the sorted cloud p64 (the gather, overlapped with the LDS sorts of other workgroups).
template <int W, int IPT>
__global__ __attribute__((amdgpu_flat_work_group_size(W * 64, W * 64), amdgpu_waves_per_eu(IPT <= 4 ? 8 : 4))) void k_seg_sort(uint64_t *ka, uint32_t *va, uint64_t *kb, uint32_t *vb,
                                                     const SegJob *__restrict__ big, const SegJob *__restrict__ small,
                                                     const uint32_t *__restrict__ njobs, uint32_t *__restrict__ next,
                                                     uint32_t *__restrict__ err, const double *__restrict__ xyz,
                                                     double4 *__restrict__ p64) {
    __shared__ SegShared<W, IPT> S;
    const uint32_t tid = threadIdx.x;
    if (njobs[2] > kSegMaxLen) return;  // a segment too long for one workgroup: the LSD passes run instead
    const uint32_t nbig = njobs[0], nsmall = njobs[1];
    for (;;) {
        if (tid == 0) {
            uint32_t j = atomicAdd(&next[0], 1u);
            if (j < nbig) {
                S.job = big[j], S.kind = 0u;
            } else {
                j = atomicAdd(&next[1], 1u);
                if (j < nsmall)
                    S.job = small[j], S.kind = 1u;
                else
                    S.kind = 2u;
            }
        }
        __syncthreads();
        const uint32_t kind = S.kind;
        const SegJob job = S.job;
        __syncthreads();
        if (kind == 2u) break;
        if (job.len > SegShared<W, IPT>::kCap)
            seg_split_sort(S, ka, va, kb, vb, job.start, job.len, err, xyz, p64);
        else
            seg_local_sort(S, ka, va, kb, vb, job.start, job.len, xyz, p64);
        __syncthreads();
    }
}

// One workgroup per top-8-bit bin: the start of each of its 16-bit segments (hist2[bin * 256 + digit]
// keys each; the bin starts at base1[bin]) and its jobs - a segment above cap alone (big), runs of
// consecutive smaller segments up to cap together (small). njobs: [0] big, [1] small, [2] longest
// segment.
__global__ __launch_bounds__(256) void k_msd_jobs(const uint32_t *__restrict__ hist2, const uint32_t *__restrict__ base1,
                                                  uint32_t cap, SegJob *__restrict__ big, SegJob *__restrict__ small,
                                                  uint32_t *__restrict__ njobs) {
    __shared__ uint32_t s_c[256], s_st[256], s_wsum[4], s_nb, s_ns, s_ob, s_os;
    __shared__ SegJob s_big[256], s_small[256];
    const uint32_t bin = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const uint32_t c = hist2[bin * 256 + tid];
    const uint32_t incl = wave_incl_scan(c);
    if (lane == 63) s_wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t ww = 0; ww < w; ww++) pre += s_wsum[ww];
    const uint32_t st = base1[bin] + pre + incl - c;
    s_c[tid] = c, s_st[tid] = st;
    __syncthreads();
    if (w == 0) {
        uint32_t c4[4], s4[4];
#pragma unroll
        for (int sl = 0; sl < 4; sl++) c4[sl] = s_c[4 * lane + sl], s4[sl] = s_st[4 * lane + sl];
        uint32_t nb = 0, ns = 0, mx = 0;  // segments up to cap never exceed the longest big one's limit
        pack_digit_ranges(
            c4, s4, cap,
            [&](uint32_t a, uint32_t l) {
                if (lane == 0) s_big[nb] = SegJob{a, l};
                nb++;
                mx = l > mx ? l : mx;
            },
            [&](uint32_t a, uint32_t l) {
                if (lane == 0) s_small[ns] = SegJob{a, l};
                ns++;
                mx = l > mx ? l : mx;
            });
        if (lane == 0) {
            s_nb = nb, s_ns = ns;
            s_ob = nb ? atomicAdd(&njobs[0], nb) : 0u;
            s_os = ns ? atomicAdd(&njobs[1], ns) : 0u;
            if (mx) atomicMax(&njobs[2], mx);
        }
    }
    __syncthreads();
    if (tid < s_nb) big[s_ob + tid] = s_big[tid];
    if (tid < s_ns) small[s_os + tid] = s_small[tid];
}

// Everything the partition passes start from, in one launch (no host copies or fills in between):
// zero look-back words, prefix counts, job and tile counters and the error flag; the bin starts of
// both passes, by block 0 from the encode kernel's histograms of bits 47..54 (h47) and 55..62 (h55).
__global__ __launch_bounds__(256) void k_part_init(const uint32_t *__restrict__ h47, const uint32_t *__restrict__ h55,
                                                   size_t nlook, uint32_t *__restrict__ lookback,
                                                   uint32_t *__restrict__ bases, uint32_t *__restrict__ hist2,
                                                   uint32_t *__restrict__ njobs, uint32_t *__restrict__ counters) {
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    for (size_t i = g; i < nlook; i += stride) lookback[i] = 0u;
    for (size_t i = g; i < 65536; i += stride) hist2[i] = 0u;
    if (g < 4) njobs[g] = 0u;
    if (g < 32) counters[g] = 0u;
    if (blockIdx.x == 0) {
        __shared__ uint32_t s_w[2][4];
        const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
        const uint32_t c0 = h47[tid], c1 = h55[tid];
        const uint32_t i0 = wave_incl_scan(c0), i1 = wave_incl_scan(c1);
        if (lane == 63) s_w[0][w] = i0, s_w[1][w] = i1;
        __syncthreads();
        uint32_t p0 = 0, p1 = 0;
        for (uint32_t ww = 0; ww < w; ww++) p0 += s_w[0][ww], p1 += s_w[1][ww];
        bases[tid] = p0 + i0 - c0, bases[256 + tid] = p1 + i1 - c1;
    }
}

// The 8 LSD digit histograms of n keys (hist must be zero), for an LSD sort after an encode that
// skipped them.
__global__ __launch_bounds__(256) void k_digit_hist(const uint64_t *__restrict__ keys, size_t n, uint32_t *__restrict__ hist) {
    __shared__ uint32_t sh[kRadixPassCount * 256];
    for (uint32_t i = threadIdx.x; i < kRadixPassCount * 256; i += 256) sh[i] = 0u;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < kRadixPassCount; p++) atomicAdd(&sh[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kRadixPassCount * 256; i += 256)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

namespace {

// LDS sort shapes: OCT_TUNE_SORT 1 = 8 waves x 4 keys (2048; four workgroups per CU, so more gather
// reads in flight), 2 = 8 waves x 8 keys (4096)
constexpr size_t kSegJobsMax = 256 * 256;  // jobs per kind: at most one per 16-bit segment

// resident workgroups of k_seg_sort<W, IPT> over the device
template <int W, int IPT>
unsigned seg_grid() {
    static unsigned cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_seg_sort<W, IPT>, W * 64, 0);
    cached = (unsigned)((cus > 0 ? cus : 256) * (per > 0 ? per : 1));
    return cached;
}

// The partition sort (above) of the context's n encoded keys, keys[0] in (left untouched, for the LSD
// passes should a 16-bit segment turn out longer than kSegMaxLen: k_seg_sort then does nothing and
// njobs[2] tells the host), keys[1] / vals[1] and the sorted cloud out. h47, h55: the histograms of
// bits 47..54 and 55..62. Launches only; returns where the njobs words are.
template <int W, int IPT>
hipError_t sort_partition(oct_ctx *ctx, uint32_t *counters, uint32_t *errflag, uint32_t **njobs_dev) {
    constexpr uint32_t cap = SegShared<W, IPT>::kCap;
    const size_t n = ctx->n;
    const uint32_t *h55 = ctx->hist.p + kRadixPassCount * 256, *h47 = h55 + 256;  // the encode kernel's
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    // msd layout (u32 words): bases [2 x 256], hist2 [65536], njobs [4], big and small jobs
    const size_t o_bases = 0, o_hist2 = 512, o_njobs = o_hist2 + 65536, o_big = o_njobs + 4, o_small = o_big + 2 * kSegJobsMax,
                 words = o_small + 2 * kSegJobsMax;
    hipError_t e;
    if ((e = ctx->msd.reserve(words)) != hipSuccess) return e;
    if ((e = ctx->lookback.reserve(2 * tiles * 256 + 1)) != hipSuccess) return e;
    if ((e = ctx->p64.reserve(n)) != hipSuccess) return e;
    if ((e = ctx->keys_p.reserve(n)) != hipSuccess) return e;
    if ((e = ctx->vals_p.reserve(n)) != hipSuccess) return e;
    uint32_t *M = ctx->msd.p;
    SegJob *big = reinterpret_cast<SegJob *>(M + o_big), *small = reinterpret_cast<SegJob *>(M + o_small);
    hipStream_t s = ctx->stream;
    const size_t nlook = 2 * tiles * 256;
    const unsigned ig = (unsigned)std::min<size_t>(2048, std::max<size_t>((nlook + 255) / 256, 256));
    k_part_init<<<ig, 256, 0, s>>>(h47, h55, nlook, ctx->lookback.p, M + o_bases, M + o_hist2, M + o_njobs, counters);
    // bits 47..54: keys[0] (values: the positions) -> keys[1] / vals[1]
    k_onesweep<false><<<(unsigned)tiles, kSortBlock, 0, s>>>(ctx->keys[0].p, nullptr, ctx->keys[1].p, ctx->vals[1].p, n,
                                                             kMsdShift2, M + o_bases, ctx->lookback.p, counters + 0, errflag,
                                                             nullptr, 0u, nullptr);
    // bits 55..62: keys[1] / vals[1] -> keys_p / vals_p, counting the keys of every 16-bit prefix
    k_onesweep<false><<<(unsigned)tiles, kSortBlock, 0, s>>>(ctx->keys[1].p, ctx->vals[1].p, ctx->keys_p.p, ctx->vals_p.p, n,
                                                             kMsdShift1, M + o_bases + 256, ctx->lookback.p + tiles * 256,
                                                             counters + 1, errflag, nullptr, 0u, M + o_hist2);
    k_msd_jobs<<<256, 256, 0, s>>>(M + o_hist2, M + o_bases + 256, cap, big, small, M + o_njobs);
    // every 16-bit segment sorted by one workgroup: keys_p / vals_p -> keys[1] / vals[1], sorted cloud
    k_seg_sort<W, IPT><<<seg_grid<W, IPT>(), W * 64, 0, s>>>(ctx->keys_p.p, ctx->vals_p.p, ctx->keys[1].p, ctx->vals[1].p,
                                                             big, small, M + o_njobs, counters + 2, errflag, ctx->xyz.p,
                                                             ctx->p64.p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    *njobs_dev = M + o_njobs;
    return hipSuccess;
}

// LSD passes (the 8-bit digits that vary) over keys[0]; h: their 8 histograms.
hipError_t sort_lsd(oct_ctx *ctx, const uint32_t *h, uint32_t *counters, uint32_t *errflag, uint32_t *bases) {
    const size_t n = ctx->n;
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
    hipError_t e;
    if ((e = ctx->lookback.reserve((size_t)(np ? np : 1) * tiles * 256 + 1)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(bases, base, sizeof(base), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(counters, 0, 16 * sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
    if (np && (e = hipMemsetAsync(ctx->lookback.p, 0, (size_t)np * tiles * 256 * sizeof(uint32_t), ctx->stream)) != hipSuccess)
        return e;
    int src = 0;
    for (int i = 0; i < np; i++) {
        int p = passes[i];
        const uint32_t *vin = (i == 0) ? nullptr : ctx->vals[src].p;
        k_onesweep<false><<<(unsigned)tiles, kSortBlock, 0, ctx->stream>>>(
            ctx->keys[src].p, vin, ctx->keys[src ^ 1].p, ctx->vals[src ^ 1].p, n, 8u * p, bases + p * 256,
            ctx->lookback.p + (size_t)i * tiles * 256, counters + i, errflag, nullptr, 0u, nullptr);
        src ^= 1;
    }
    if (np == 0) {
        int nb = (int)((n + 255) / 256);
        nb = nb < 1 ? 1 : (nb > 4096 ? 4096 : nb);
        k_iota_copy<<<nb, 256, 0, ctx->stream>>>(ctx->keys[0].p, ctx->keys[1].p, ctx->vals[1].p, n);
        src = 1;
    }
    ctx->sorted_keys = ctx->keys[src].p;
    ctx->perm = ctx->vals[src].p;
    ctx->p64_sorted = false;
    return hipGetLastError();
}

// OCT_TUNE_SORT: 0 = LSD passes, 1 / 2 = partition sort with 2048- / 4096-key LDS sorts (1 by default)
// the default form is the partition sort, unless a sort of the current points already had to fall
// back to the LSD passes (a 16-bit segment above the workgroup limit): then the LSD passes directly
int sort_mode(const oct_ctx *ctx) {
    if (ctx->sort_override >= 0) return ctx->sort_override;
    return ctx->sort_lsd_learned ? 0 : 1;
}

}  // namespace

bool sort_partition_planned(const oct_ctx *ctx) {
    const int mode = sort_mode(ctx);
    return mode > 0 && ctx->n > (mode == 2 ? 4096u : 2048u);
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
    // misc layout: [0..16) tile counters, [16] error flag, [32..32+2048) bin bases
    if ((e = ctx->misc.reserve(12 * 1024 + 64)) != hipSuccess) return e;
    uint32_t *counters = ctx->misc.p, *errflag = ctx->misc.p + 16, *bases = ctx->misc.p + 32;
    ctx->p64_sorted = false;
    // keys come from the 3-D encoders (63 bits), so bits 47..62 are the top 16 the partition sort uses
    const int mode = sort_mode(ctx);
    const bool partition = sort_partition_planned(ctx);
    // the LSD passes are planned on the host from the encode kernel's histograms; the partition sort
    // takes its bin starts from them on the device, so nothing is read back before it
    uint32_t h[kEncHists * 256];
    if (!partition) {
        if ((e = hipMemcpyAsync(h, ctx->hist.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    }
    stage_begin(ctx, ST_SORT);
    uint32_t *njobs_dev = nullptr;
    if (partition) {
        e = mode == 2 ? sort_partition<8, 8>(ctx, counters, errflag, &njobs_dev)
                      : sort_partition<8, 4>(ctx, counters, errflag, &njobs_dev);
        if (e != hipSuccess) return e;
        ctx->p64_sorted = true;
        ctx->sorted_keys = ctx->keys[1].p;
        ctx->perm = ctx->vals[1].p;
    } else {
        if ((e = hipMemsetAsync(counters, 0, 32 * sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
        if (!ctx->hist_lsd) {  // the encode kernel wrote only the partition histogram
            if ((e = hipMemsetAsync(ctx->hist.p, 0, kRadixPassCount * 256 * sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
            const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1024));
            k_digit_hist<<<nb, 256, 0, ctx->stream>>>(ctx->keys[0].p, n, ctx->hist.p);
            if ((e = hipMemcpyAsync(h, ctx->hist.p, kRadixPassCount * 256 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    ctx->stream)) != hipSuccess)
                return e;
            if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
            ctx->hist_lsd = true;
        }
        if ((e = sort_lsd(ctx, h, counters, errflag, bases)) != hipSuccess) return e;
    }
    stage_end(ctx, ST_SORT);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t fb[2] = {0, 0};  // error flag, longest 16-bit segment
    if ((e = hipMemcpyAsync(&fb[0], errflag, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
    if (njobs_dev && (e = hipMemcpyAsync(&fb[1], njobs_dev + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    if (!fb[0] && partition && fb[1] > kSegMaxLen) {
        // a segment too long for one workgroup: the LSD passes over the untouched encoded keys, timed
        // as part of the same stage; later default-form sorts of these points go to them directly
        ctx->sort_lsd_learned = true;
        if (!ctx->hist_lsd) {
            if ((e = hipMemsetAsync(ctx->hist.p, 0, kRadixPassCount * 256 * sizeof(uint32_t), ctx->stream)) != hipSuccess) return e;
            const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1024));
            k_digit_hist<<<nb, 256, 0, ctx->stream>>>(ctx->keys[0].p, n, ctx->hist.p);
            if ((e = hipMemcpyAsync(h, ctx->hist.p, kRadixPassCount * 256 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    ctx->stream)) != hipSuccess)
                return e;
            if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
            ctx->hist_lsd = true;
        }
        if ((e = sort_lsd(ctx, h, counters, errflag, bases)) != hipSuccess) return e;
        stage_end(ctx, ST_SORT);
        if ((e = hipMemcpyAsync(&fb[0], errflag, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    }
    if (fb[0]) {
        err = (fb[0] & 2u) ? "sort: a segment split overflowed its stack" : "sort: look-back wait exceeded its bound";
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
                                                                counters + p, err, tiles, kSegBaseStride, nullptr);
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
                                                                 counters + p, err, nullptr, 0u, nullptr);
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
    if (ctx->n && !ctx->p64_sorted) {  // the partition sort writes the sorted cloud itself
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
