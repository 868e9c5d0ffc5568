// Device-side helpers: quantisation, Morton / Hilbert keys, octree key algebra, wave64 primitives.
// All FP64 arithmetic here is built with -ffp-contract=off so it rounds exactly as the reference's
// x86 build (no FMA, cmake/CMakeConfig.cmake:41,44).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace oct {

// ------------------------------------------------------------------ quantisation
// The reference converts with (coords_t)(x_t * 2^21) where coords_t is a 64-bit unsigned type
// (point_encoder.hpp:17, 46-48). On its x86-64 build (no AVX-512) g++ lowers that conversion to
// cvttsd2si for d < 2^63 (NaN or out of int64 range -> INT64_MIN) and to cvttsd2si(d - 2^63) ^ 2^63
// above. After min(., 2^21 - 1) that gives: NaN -> 2^21-1, d <= -1 -> 2^21-1, (-1, 2^21) -> trunc(d),
// [2^21, 2^64) -> 2^21-1, >= 2^64 (incl. +inf) -> 0.
__host__ __device__ inline uint32_t clamp_coord(double d) {
    if (!(d > -1.0)) return kMaxCoord;                        // NaN, d <= -1, -inf
    if (d < 2097152.0) return (uint32_t)(d < 0.0 ? 0.0 : d);  // (-1, 2^21): truncation toward zero
    if (d >= 18446744073709551616.0) return 0u;               // >= 2^64
    return kMaxCoord;
}

// getAnchorCoords (point_encoder.hpp:37-49), one axis: ((p - c) + r) / (2 * r) * 2^21.
__host__ __device__ inline uint32_t quantize_axis(double p, double c, double r) {
    double t = ((p - c) + r) / (2.0 * r);
    return clamp_coord(t * 2097152.0);
}

// ------------------------------------------------------------------ Morton (Z-order)
// Bit j of x goes to key bit 3j, y to 3j+1, z to 3j+2 (libmorton/morton_BMI.h:40-51). Spread by
// shift-and-mask doubling.
__host__ __device__ inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x001f00000000ffffull;
    v = (v | (v << 16)) & 0x001f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
__host__ __device__ inline uint32_t compact3(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x001f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x001f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}
__host__ __device__ inline uint64_t morton_encode(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__host__ __device__ inline void morton_decode(uint64_t k, uint32_t &x, uint32_t &y, uint32_t &z) {
    x = compact3(k), y = compact3(k >> 1), z = compact3(k >> 2);
}

// ------------------------------------------------------------------ Hilbert
// The reference's encoder (hilbert_encoder_3d.hpp:29-66) keeps three working registers that are,
// at every level, the original coordinates under an axis permutation and per-axis complement.
// That (permutation, complement) pair is a 48-value state, so the per-level update becomes a
// table lookup: table[state][original octant bits] = digit | next_state << 3. The table is derived
// at compile time by running the reference's update rules on the symbolic state.
struct HilbertTable {
    uint16_t t[48 * 8];
};
constexpr int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int perm_index(int a, int b, int c) {
    for (int i = 0; i < 6; i++)
        if (kPerm[i][0] == a && kPerm[i][1] == b && kPerm[i][2] == c) return i;
    return -1;
}
constexpr HilbertTable make_hilbert_table() {
    HilbertTable h{};
    const uint32_t m2h[8] = {0, 1, 3, 2, 7, 6, 4, 5};  // hilbert_encoder_3d.hpp:22
    for (int s = 0; s < 48; s++) {
        int p[3] = {kPerm[s / 8][0], kPerm[s / 8][1], kPerm[s / 8][2]};
        int f[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
        for (int o = 0; o < 8; o++) {
            int b[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};  // original x, y, z bits
            uint32_t xi = (uint32_t)(b[p[0]] ^ f[0]), yi = (uint32_t)(b[p[1]] ^ f[1]), zi = (uint32_t)(b[p[2]] ^ f[2]);
            uint32_t digit = m2h[(xi << 2) | (yi << 1) | zi];
            int nf[3] = {f[0] ^ (int)(xi & ((!yi) | zi)), f[1] ^ (int)((xi & (yi | zi)) | (yi & (!zi))),
                         f[2] ^ (int)((xi & (!yi) & (!zi)) | (yi & (!zi)))};
            int np[3] = {p[0], p[1], p[2]}, nf2[3] = {nf[0], nf[1], nf[2]};
            if (zi) {  // x, y, z <- y, z, x
                np[0] = p[1], np[1] = p[2], np[2] = p[0];
                nf2[0] = nf[1], nf2[1] = nf[2], nf2[2] = nf[0];
            } else if (!yi) {  // swap x and z
                np[0] = p[2], np[2] = p[0];
                nf2[0] = nf[2], nf2[2] = nf[0];
            }
            int ns = perm_index(np[0], np[1], np[2]) * 8 + (nf2[0] << 2) + (nf2[1] << 1) + nf2[2];
            h.t[s * 8 + o] = (uint16_t)(digit | (uint32_t)ns << 3);
        }
    }
    return h;
}
constexpr HilbertTable kHilbert = make_hilbert_table();
constexpr int kHilbertStart = 0;  // identity permutation, no complement

// Encode with the state table held in `tab` (LDS or constant memory).
__host__ __device__ inline uint64_t hilbert_encode_tab(const uint16_t *tab, uint32_t x, uint32_t y, uint32_t z) {
    uint64_t key = 0;
    uint32_t s = kHilbertStart;
    for (int level = (int)kMaxDepth - 1; level >= 0; level--) {
        uint32_t o = (((x >> level) & 1u) << 2) | (((y >> level) & 1u) << 1) | ((z >> level) & 1u);
        uint32_t e = tab[s * 8 + o];
        key = (key << 3) | (e & 7u);
        s = e >> 3;
    }
    return key;
}

// Two levels per lookup: tab2[state * 64 + (upper level's octant << 3 | lower level's octant)] =
// the two digits (upper << 3 | lower) | the state after both << 6, composed from the one-level table.
constexpr int kHilbertTab2 = 48 * 64;
__host__ __device__ inline void hilbert_tab2_fill(uint16_t *tab2, const uint16_t *tab, uint32_t first, uint32_t step) {
    for (uint32_t i = first; i < (uint32_t)kHilbertTab2; i += step) {
        const uint32_t s = i >> 6, oh = (i >> 3) & 7u, ol = i & 7u;
        const uint32_t e1 = tab[s * 8 + oh], e2 = tab[(e1 >> 3) * 8 + ol];
        tab2[i] = (uint16_t)(((e1 & 7u) << 3) | (e2 & 7u) | ((e2 >> 3) << 6));
    }
}
// Every level's octant in one word: bits [3L, 3L + 3) = x bit L << 2 | y bit L << 1 | z bit L.
__host__ __device__ inline uint64_t hilbert_octants(uint32_t x, uint32_t y, uint32_t z) {
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}
// The key of hilbert_encode_tab from the octant word: the top level by the one-level table, the 20
// levels below it two at a time (11 dependent lookups instead of 21, and no per-level bit gathering).
__host__ __device__ inline uint64_t hilbert_encode_tab2(const uint16_t *tab, const uint16_t *tab2, uint32_t x, uint32_t y,
                                                        uint32_t z) {
    const uint64_t m = hilbert_octants(x, y, z);
    const uint32_t e = tab[kHilbertStart * 8 + (uint32_t)(m >> (3 * (kMaxDepth - 1)))];
    uint64_t key = e & 7u;
    uint32_t s = e >> 3;
#pragma unroll
    for (int level = (int)kMaxDepth - 3; level >= 0; level -= 2) {  // levels (level + 1, level)
        const uint32_t e2 = tab2[s * 64 + ((uint32_t)(m >> (3 * level)) & 63u)];
        key = (key << 6) | (e2 & 63u);
        s = e2 >> 6;
    }
    return key;
}

// HilbertEncoder3D::decode (hilbert_encoder_3d.hpp:256-289); used per node, not per point.
__host__ __device__ inline void hilbert_decode(uint64_t code, uint32_t &ox, uint32_t &oy, uint32_t &oz) {
    uint32_t x = 0, y = 0, z = 0;
    for (uint32_t level = 0; level < kMaxDepth; level++) {
        const uint32_t octant = (uint32_t)(code >> (3 * level)) & 7u;
        const uint32_t xi = octant >> 2u, yi = (octant >> 1u) & 1u, zi = octant & 1u;
        if (yi ^ zi) {
            uint32_t t = x;
            x = z, z = y, y = t;
        } else if ((!xi & !yi & !zi) || (xi & yi & zi)) {
            uint32_t t = x;
            x = z, z = t;
        }
        const uint32_t mask = (1u << level) - 1u;
        x ^= mask & (0u - (xi & (yi | zi)));
        y ^= mask & (0u - ((xi & ((!yi) | (!zi))) | ((!xi) & yi & zi)));
        z ^= mask & (0u - ((xi & (!yi) & (!zi)) | (yi & zi)));
        x |= (xi << level);
        y |= ((xi ^ yi) << level);
        z |= ((yi ^ zi) << level);
    }
    ox = x, oy = y, oz = z;
}

// ------------------------------------------------------------------ key algebra (point_encoder.hpp:542-645)
__host__ __device__ inline uint64_t node_range(uint32_t level) { return 1ull << (3u * (kMaxDepth - level)); }
__host__ __device__ inline uint64_t encode_placeholder(uint64_t key, uint32_t level) {
    return (1ull << (3u * level)) | (key >> (3u * (kMaxDepth - level)));
}

// getCenter (point_encoder.hpp:517-539), evaluated in the reference's order.
__host__ __device__ inline void node_center(int encoder, uint64_t key, uint32_t level, const double boxmin[3],
                                            const double hl[3], const double rad[3], double out[3]) {
    uint32_t c[3];
    if (encoder == OCT_ENC_HILBERT3D)
        hilbert_decode(key, c[0], c[1], c[2]);
    else
        morton_decode(key, c[0], c[1], c[2]);
    const uint32_t mask = ((1u << kMaxDepth) - 1u) ^ ((1u << (kMaxDepth - level)) - 1u);
    for (int a = 0; a < 3; a++) {
        double m = (double)(c[a] & mask);
        double v = boxmin[a] + m * hl[a] * 2.0;
        out[a] = v + rad[a];
    }
}

// ------------------------------------------------------------------ wave64 helpers
__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }
__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}
// v of lane (lane ^ OFF) for OFF < 32 (the xor mode of ds_swizzle within 32-lane halves: no lane-index
// arithmetic), and of lane (lane ^ 32) with a ds_bpermute at the address (lane ^ 32) * 4. Every lane
// of the wave active.
template <int OFF>
__device__ inline uint32_t lane_xor(uint32_t v) {
    static_assert(OFF > 0 && OFF < 32 && (OFF & (OFF - 1)) == 0, "a power of two below 32");
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (OFF << 10) | 0x1f);
}
__device__ inline uint32_t lane_xor32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane_id() ^ 32u) << 2), (int)v);
}
__device__ inline uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint32_t readlane_u32(uint32_t v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ inline double readlane_f64(double v, int lane) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Device-scope relaxed atomic accesses for self-contained flag words (flag and value in one 32-bit
// word, nothing else published through them). They are coherent across the XCDs' L2s without the
// whole-L2 writeback / invalidate that release / acquire ordering adds to every access.
__device__ inline uint32_t load_flag(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_flag(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace oct
