/*
 * oracle.c -- CPU restatement of the reference hot path. TEST INFRASTRUCTURE ONLY (see oracle.h):
 * the product never links this file; tests, smoke() and bench.py's cpu_baseline leg load it as the
 * checker or as the timed CPU baseline.
 *
 * Built with -O2 -fopenmp -ffp-contract=off (oracle/Makefile). Citations are file:line in the
 * reference repository (derivada/octrees-benchmark).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAX_DEPTH 21u
#define UPPER_BOUND 0x8000000000000000ull /* morton_encoder_3d.hpp:29, hilbert_encoder_3d.hpp:16 */
#define UNUSED_BITS 1u

int orc_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ geometry */

/* mbb (box.hpp:94-137): per-axis min/max with std::min/std::max semantics, then
 * center = min + (max - min) / 2 (util.hpp:139-143) and radii = (max - min) / 2 (box.hpp:19). */
void orc_mbb(const double *xyz, size_t n, double center[3], double radii[3]) {
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
#pragma omp parallel
    {
        double lmn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
        double lmx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
#pragma omp for schedule(static) nowait
        for (size_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                double v = xyz[3 * i + a];
                lmn[a] = (v < lmn[a]) ? v : lmn[a]; /* std::min(a, b) == (b < a) ? b : a */
                lmx[a] = (lmx[a] < v) ? v : lmx[a]; /* std::max(a, b) == (a < b) ? b : a */
            }
        }
#pragma omp critical
        {
            for (int a = 0; a < 3; ++a) {
                mn[a] = (lmn[a] < mn[a]) ? lmn[a] : mn[a];
                mx[a] = (mx[a] < lmx[a]) ? lmx[a] : mx[a];
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        center[a] = mn[a] + (mx[a] - mn[a]) / 2;
        radii[a] = (mx[a] - mn[a]) / 2;
    }
}

/* x86-64 g++ lowering of (uint64_t)d without AVX-512 (observed with g++ -O2 -mavx2 -mbmi2):
 * d < 2^63 or NaN -> cvttsd2si (NaN / out of range -> INT64_MIN); else cvttsd2si(d - 2^63) ^ 2^63. */
static uint64_t x86_cvt_u64(double d) {
    const double two63 = 9223372036854775808.0;
    if (!(d >= two63)) {
        if (d != d || d <= -two63 - 1.0 || d >= two63) return 0x8000000000000000ull;
        return (uint64_t)(int64_t)d;
    }
    double v = d - two63;
    uint64_t r = (v >= two63) ? 0x8000000000000000ull : (uint64_t)(int64_t)v;
    return r ^ 0x8000000000000000ull;
}

/* getAnchorCoords (point_encoder.hpp:37-49), one axis:
 * x_t = ((p - c) + r) / (2 * r); x = min((coords_t)(x_t * 2^21), 2^21 - 1). */
uint64_t orc_quantize_axis(double p, double c, double r) {
    double t = ((p - c) + r) / (2 * r);
    uint64_t q = x86_cvt_u64(t * (double)(1 << MAX_DEPTH));
    uint64_t mx = (1u << MAX_DEPTH) - 1u;
    return q < mx ? q : mx;
}

/* libmorton m3D_e_BMI / m3D_d_BMI (libmorton/morton_BMI.h:40-58): pdep/pext with masks
 * X = 0x9249..., Y = 0x2492..., Z = 0x4924...; written as a portable bit loop. */
uint64_t orc_morton_encode(uint64_t x, uint64_t y, uint64_t z) {
    uint64_t k = 0;
    for (unsigned j = 0; j < 22; ++j) {
        if (3 * j < 64) k |= ((x >> j) & 1ull) << (3 * j);
        if (3 * j + 1 < 64) k |= ((y >> j) & 1ull) << (3 * j + 1);
        if (3 * j + 2 < 64) k |= ((z >> j) & 1ull) << (3 * j + 2);
    }
    return k;
}

void orc_morton_decode(uint64_t key, uint64_t out[3]) {
    uint64_t x = 0, y = 0, z = 0;
    for (unsigned j = 0; j < 22; ++j) {
        if (3 * j < 64) x |= ((key >> (3 * j)) & 1ull) << j;
        if (3 * j + 1 < 64) y |= ((key >> (3 * j + 1)) & 1ull) << j;
        if (3 * j + 2 < 64) z |= ((key >> (3 * j + 2)) & 1ull) << j;
    }
    out[0] = x, out[1] = y, out[2] = z;
}

/* HilbertEncoder3D::encode (hilbert_encoder_3d.hpp:22, 29-66). 32-bit lane arithmetic is exact for
 * the 21 used bits: the reference's `x ^= -(uint32 ...)` flips only the low 32 bits and the
 * rotations truncate to uint32_t. */
uint64_t orc_hilbert_encode(uint64_t x64, uint64_t y64, uint64_t z64) {
    static const uint32_t m2h[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    uint32_t x = (uint32_t)x64, y = (uint32_t)y64, z = (uint32_t)z64;
    uint64_t key = 0;
    for (int level = MAX_DEPTH - 1; level >= 0; level--) {
        const uint32_t xi = (x >> level) & 1u;
        const uint32_t yi = (y >> level) & 1u;
        const uint32_t zi = (z >> level) & 1u;
        const uint32_t octant = (xi << 2) | (yi << 1) | zi;
        key <<= 3;
        key |= m2h[octant];
        x ^= -(xi & ((!yi) | zi));
        y ^= -((xi & (yi | zi)) | (yi & (!zi)));
        z ^= -((xi & (!yi) & (!zi)) | (yi & (!zi)));
        if (zi) {
            uint32_t t = x;
            x = y, y = z, z = t;
        } else if (!yi) {
            uint32_t t = x;
            x = z, z = t;
        }
    }
    return key;
}

/* HilbertEncoder3D::decode (hilbert_encoder_3d.hpp:256-289). */
void orc_hilbert_decode(uint64_t code, uint64_t out[3]) {
    uint64_t x = 0, y = 0, z = 0;
    for (unsigned level = 0; level < MAX_DEPTH; level++) {
        const uint64_t octant = (code >> (3 * level)) & 7u;
        const uint64_t xi = octant >> 2u;
        const uint64_t yi = (octant >> 1u) & 1u;
        const uint64_t zi = octant & 1u;
        if (yi ^ zi) {
            uint64_t t = x;
            x = z, z = y, y = t;
        } else if ((!xi & !yi & !zi) || (xi & yi & zi)) {
            uint64_t t = x;
            x = z, z = t;
        }
        const uint64_t mask = (1u << level) - 1u;
        x ^= mask & (-(xi & (yi | zi)));
        y ^= mask & (-((xi & ((!yi) | (!zi))) | ((!xi) & yi & zi)));
        z ^= mask & (-((xi & (!yi) & (!zi)) | (yi & zi)));
        x |= (xi << level);
        y |= ((xi ^ yi) << level);
        z |= ((yi ^ zi) << level);
    }
    out[0] = x, out[1] = y, out[2] = z;
}

static void decode_key(int enc, uint64_t key, uint64_t out[3]) {
    if (enc == ORC_ENC_HILBERT3D)
        orc_hilbert_decode(key, out);
    else
        orc_morton_decode(key, out);
}

/* encodePoints (point_encoder.hpp:54-63) -> encodeFromPoint -> getAnchorCoords + encode. */
void orc_encode(int enc, const double *xyz, size_t n, const double center[3], const double radii[3],
                uint64_t *keys) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        uint64_t x = orc_quantize_axis(xyz[3 * i + 0], center[0], radii[0]);
        uint64_t y = orc_quantize_axis(xyz[3 * i + 1], center[1], radii[1]);
        uint64_t z = orc_quantize_axis(xyz[3 * i + 2], center[2], radii[2]);
        keys[i] = (enc == ORC_ENC_HILBERT3D) ? orc_hilbert_encode(x, y, z) : orc_morton_encode(x, y, z);
    }
}

/* sortPoints radix passes (point_encoder.hpp:293-376): 8 passes x 8 bits; per-thread histograms
 * over static contiguous chunks; bucket-major, thread-minor offsets; stable scatter. The chunks run
 * on `nthreads` OpenMP threads, as the reference's; the result does not depend on it. */
void orc_sort(uint64_t *keys, uint32_t *perm, size_t n, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    uint64_t *kb = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint32_t *pb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t *hist = (uint64_t *)calloc((size_t)nthreads * 256, sizeof(uint64_t));
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    uint64_t *ks = keys, *kd = kb;
    uint32_t *ps = perm, *pd = pb;
    for (int pass = 0; pass < 8; pass++) {
        int shift = pass * 8;
        memset(hist, 0, (size_t)nthreads * 256 * sizeof(uint64_t));
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
        for (int t = 0; t < nthreads; t++) {
            size_t lo = n * (size_t)t / (size_t)nthreads, hi = n * (size_t)(t + 1) / (size_t)nthreads;
            for (size_t i = lo; i < hi; i++) hist[t * 256 + ((ks[i] >> shift) & 255)]++;
        }
        uint64_t off = 0;
        for (int b = 0; b < 256; b++)
            for (int t = 0; t < nthreads; t++) {
                uint64_t v = hist[t * 256 + b];
                hist[t * 256 + b] = off;
                off += v;
            }
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
        for (int t = 0; t < nthreads; t++) {
            size_t lo = n * (size_t)t / (size_t)nthreads, hi = n * (size_t)(t + 1) / (size_t)nthreads;
            for (size_t i = lo; i < hi; i++) {
                uint64_t pos = hist[t * 256 + ((ks[i] >> shift) & 255)]++;
                kd[pos] = ks[i];
                pd[pos] = ps[i];
            }
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint32_t *tp = ps; ps = pd; pd = tp;
    }
    /* 8 passes: data ends back in the caller's arrays */
    free(kb);
    free(pb);
    free(hist);
}

/* The points reordered by the sort's permutation (sortPoints moves the Points with their keys,
 * point_encoder.hpp:359-375): out[i] = xyz[perm[i]], OpenMP over the output. */
void orc_gather3(const double *xyz, const uint32_t *perm, size_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        const double *p = &xyz[3 * (size_t)perm[i]];
        out[3 * i] = p[0], out[3 * i + 1] = p[1], out[3 * i + 2] = p[2];
    }
}

/* ------------------------------------------------------------------ key algebra
 * point_encoder.hpp:542-645 */
static inline uint32_t clz64(uint64_t x) { return x ? (uint32_t)__builtin_clzll(x) : 64u; }
static inline uint32_t get_level(uint64_t range) {
    if (range == UPPER_BOUND) return 0;
    return (clz64(range - 1) - UNUSED_BITS) / 3;
}
static inline uint32_t sibling_id(uint64_t code, uint32_t level) {
    return (uint32_t)((code >> (3ull * (MAX_DEPTH - level))) & 7ull);
}
static inline uint64_t node_range(uint32_t level) { return 1ull << (3ull * (MAX_DEPTH - level)); }
static inline uint32_t decode_prefix_length(uint64_t code) { return 63u - clz64(code); }
static inline uint64_t encode_placeholder(uint64_t code, uint32_t level) {
    uint64_t ret = code >> (3 * (MAX_DEPTH - level));
    return (1ull << (3 * level)) | ret;
}
static inline uint64_t decode_placeholder(uint64_t code) {
    uint32_t pl = decode_prefix_length(code);
    uint64_t ret = code ^ (1ull << pl);
    return ret << (3ull * MAX_DEPTH - pl);
}
static inline int32_t common_prefix(uint64_t a, uint64_t b) { return (int32_t)clz64(a ^ b) - (int32_t)UNUSED_BITS; }
static inline uint32_t octal_digit(uint64_t code, uint32_t position) {
    return (uint32_t)((code >> (3ull * (MAX_DEPTH - position))) & 7ull);
}
static inline uint32_t log8ceil(uint64_t n) {
    if (n == 0) return 0;
    uint32_t lz = clz64(n - 1);
    return MAX_DEPTH - (lz - UNUSED_BITS) / 3;
}

/* ------------------------------------------------------------------ octree */
struct orc_tree {
    uint32_t nLeaf, nInternal, nTotal, maxDepthSeen, rounds;
    uint64_t *leaves, *counts, *layout;
    uint64_t *prefixes;
    uint32_t *offsets;
    uint64_t *ranges; /* 2 per node */
    double *centers;  /* 3 per node */
    double radii[22 * 3];
    double boxmin[3], boxmax[3];
    double halfLengths[3];
    int enc;
};

static size_t lower_bound_u64(const uint64_t *a, size_t lo, size_t hi, uint64_t v) {
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
static size_t upper_bound_u64(const uint64_t *a, size_t lo, size_t hi, uint64_t v) {
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* calculateNodeOp + siblingAndLevel (linear_octree.hpp:173-222) */
static uint64_t calc_node_op(const uint64_t *leaves, const uint64_t *counts, uint32_t index,
                             uint32_t maxLeaf, uint32_t *maxDepthSeen) {
    uint64_t node = leaves[index];
    uint64_t range = leaves[index + 1] - node;
    uint32_t level = get_level(range);
    int32_t sibling;
    if (level == 0) {
        sibling = -1;
    } else {
        uint32_t sid = sibling_id(node, level);
        int on_tree = leaves[index - sid + 8] == (leaves[index - sid] + node_range(level - 1));
        sibling = on_tree ? (int32_t)sid : -1;
    }
    if (sibling > 0) {
        uint32_t p = index - (uint32_t)sibling;
        uint64_t parentCount = counts[p] + counts[p + 1] + counts[p + 2] + counts[p + 3] + counts[p + 4] +
                               counts[p + 5] + counts[p + 6] + counts[p + 7];
        if (parentCount <= maxLeaf) return 0;
    }
    uint32_t nodeCount = (uint32_t)counts[index];
    if (nodeCount > maxLeaf && level < MAX_DEPTH) {
        if (level + 1 > *maxDepthSeen) *maxDepthSeen = level + 1;
        return 8;
    }
    return 1;
}

/* computeNodeCounts (linear_octree.hpp:291-321) */
static void compute_node_counts(const uint64_t *leaves, size_t nl, uint64_t *counts, const uint64_t *codes,
                                size_t n) {
    size_t firstNode = 0, lastNode = nl;
    if (n > 0) {
        firstNode = upper_bound_u64(leaves, 0, nl + 1, codes[0]) - 1;
        lastNode = upper_bound_u64(leaves, 0, nl + 1, codes[n - 1]);
    } else {
        firstNode = nl, lastNode = nl;
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nl; i++) {
        if (i < firstNode || i >= lastNode)
            counts[i] = 0;
        else
            counts[i] = lower_bound_u64(codes, 0, n, leaves[i + 1]) - lower_bound_u64(codes, 0, n, leaves[i]);
    }
}

static int32_t digit_weight(uint32_t digit) {
    int32_t m = -(int32_t)(digit >= 4);
    return ((7 - (int32_t)digit) & m) - ((int32_t)digit & ~m);
}
static int32_t binary_key_weight(uint64_t key, uint32_t level) {
    int32_t ret = 0;
    for (uint32_t l = 1; l <= level + 1; ++l) ret += digit_weight(octal_digit(key, l));
    return ret;
}

typedef struct {
    uint64_t key;
    int32_t val;
} kv_t;

/* stable merge sort by key (std::stable_sort at linear_octree.hpp:517-519) */
static void stable_sort_kv(kv_t *a, size_t n) {
    kv_t *tmp = (kv_t *)malloc((n ? n : 1) * sizeof(kv_t));
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = (a[j].key < a[i].key) ? a[j++] : a[i++];
            while (i < mid) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, n * sizeof(kv_t));
    }
    free(tmp);
}

/* computeInternalRanges (linear_octree.hpp:410-428), post-order recursion */
static void internal_ranges(orc_tree *t, const int32_t *internalToLeaf, uint32_t node) {
    if (t->offsets[node] == 0) {
        int32_t li = internalToLeaf[node];
        t->ranges[2 * node] = t->layout[li];
        t->ranges[2 * node + 1] = t->layout[li + 1];
        return;
    }
    for (uint32_t oct = 0; oct < 8; oct++) {
        uint32_t child = t->offsets[node] + oct;
        internal_ranges(t, internalToLeaf, child);
        if (oct == 0) t->ranges[2 * node] = t->ranges[2 * child];
        if (oct == 7) t->ranges[2 * node + 1] = t->ranges[2 * child + 1];
    }
}

/* getCenter (point_encoder.hpp:517-539) */
static void get_center(const orc_tree *t, uint64_t code, uint32_t level, double out[3]) {
    uint64_t c[3];
    decode_key(t->enc, code, c);
    uint64_t mask = ((1u << MAX_DEPTH) - 1) ^ ((1u << (MAX_DEPTH - level)) - 1);
    for (int a = 0; a < 3; a++) {
        double m = (double)(c[a] & mask);
        double v = t->boxmin[a] + m * t->halfLengths[a] * 2;
        out[a] = v + t->radii[3 * level + a];
    }
}

/* LinearOctree ctor -> buildOctree (linear_octree.hpp:567-623) */
orc_tree *orc_build(const uint64_t *codes, size_t n, uint32_t maxLeaf, int enc, const double center[3],
                    const double radii[3]) {
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->enc = enc;
    /* Box(center, radii): min = center - radii, max = center + radii (box.hpp:17) */
    for (int a = 0; a < 3; a++) {
        t->boxmin[a] = center[a] - radii[a];
        t->boxmax[a] = center[a] + radii[a];
    }
    /* setupBbox (linear_octree.hpp:439-455); eps = 1.0f / (1 << 21) */
    const double eps = (double)(1.0f / (float)(1 << MAX_DEPTH));
    for (int a = 0; a < 3; a++) t->halfLengths[a] = (double)0.5f * eps * (t->boxmax[a] - t->boxmin[a]);
    for (uint32_t i = 0; i <= MAX_DEPTH; i++) {
        uint64_t side = 1u << (MAX_DEPTH - i);
        for (int a = 0; a < 3; a++) t->radii[3 * i + a] = (double)side * t->halfLengths[a];
    }

    /* buildOctreeLeaves (linear_octree.hpp:458-489) */
    size_t nl = 1;
    uint64_t *leaves = (uint64_t *)malloc(2 * sizeof(uint64_t));
    uint64_t *counts = (uint64_t *)malloc(1 * sizeof(uint64_t));
    leaves[0] = 0, leaves[1] = UPPER_BOUND;
    counts[0] = n;
    t->rounds = 0;
    for (;;) {
        uint64_t *ops = (uint64_t *)calloc(nl + 1, sizeof(uint64_t));
        int converged = 1;
        uint32_t mds = t->maxDepthSeen;
        for (size_t i = 0; i < nl; i++) {
            ops[i] = calc_node_op(leaves, counts, (uint32_t)i, maxLeaf, &mds);
            if (ops[i] != 1) converged = 0;
        }
        t->maxDepthSeen = mds;
        t->rounds++;
        if (converged) {
            free(ops);
            break;
        }
        /* rebalanceTree: exclusive scan then processNode (linear_octree.hpp:232-283, 325-333) */
        uint64_t a = 0, b = 0;
        for (size_t i = 0; i < nl + 1; i++) {
            a += ops[i];
            ops[i] = b;
            b = a;
        }
        size_t nnl = ops[nl];
        uint64_t *nleaves = (uint64_t *)malloc((nnl + 1) * sizeof(uint64_t));
        nleaves[nnl] = leaves[nl];
        for (size_t i = 0; i < nl; i++) {
            uint64_t opCode = ops[i + 1] - ops[i];
            if (opCode == 0) continue;
            uint64_t node = leaves[i];
            uint32_t level = get_level(leaves[i + 1] - node);
            uint64_t ni = ops[i];
            nleaves[ni] = node;
            if (opCode > 1) {
                uint32_t levelDiff = log8ceil(opCode);
                uint64_t gap = node_range(level + levelDiff);
                for (uint64_t s = 1; s < opCode; s++) nleaves[ni + s] = nleaves[ni + s - 1] + gap;
            }
        }
        free(ops);
        free(leaves);
        free(counts);
        leaves = nleaves;
        nl = nnl;
        counts = (uint64_t *)malloc(nl * sizeof(uint64_t));
        compute_node_counts(leaves, nl, counts, codes, n);
    }
    t->nLeaf = (uint32_t)nl;
    t->leaves = leaves;
    t->counts = counts;
    t->layout = (uint64_t *)malloc((nl + 1) * sizeof(uint64_t));
    t->layout[0] = 0;
    for (size_t i = 0; i < nl; i++) t->layout[i + 1] = t->layout[i] + counts[i];

    /* resize (linear_octree.hpp:491-503) */
    uint32_t nInternal = (t->nLeaf - 1) / 7;
    uint32_t nTotal = t->nLeaf + nInternal;
    t->nInternal = nInternal, t->nTotal = nTotal;
    uint64_t *prefixes = (uint64_t *)calloc(nTotal, sizeof(uint64_t));
    int32_t *internalToLeaf = (int32_t *)calloc(nTotal, sizeof(int32_t));
    int32_t *leafToInternal = (int32_t *)calloc(nTotal, sizeof(int32_t));
    t->offsets = (uint32_t *)calloc((size_t)nTotal + 1, sizeof(uint32_t));
    t->ranges = (uint64_t *)calloc(2 * (size_t)nTotal, sizeof(uint64_t));
    t->centers = (double *)calloc(3 * (size_t)nTotal, sizeof(double));

    /* createUnsortedLayout (linear_octree.hpp:353-367) */
    for (size_t i = 0; i < nl; i++) {
        uint64_t key = leaves[i];
        uint32_t level = get_level(leaves[i + 1] - key);
        prefixes[i + nInternal] = encode_placeholder(key, level);
        internalToLeaf[i + nInternal] = (int32_t)(i + nInternal);
        uint32_t prefixLength = (uint32_t)common_prefix(key, leaves[i + 1]);
        if (prefixLength % 3 == 0 && i < nl - 1) {
            uint32_t octIndex = (uint32_t)((i + (size_t)(int64_t)binary_key_weight(key, prefixLength / 3)) / 7);
            if (octIndex >= nTotal) goto fail;
            prefixes[octIndex] = encode_placeholder(key, prefixLength / 3);
            internalToLeaf[octIndex] = (int32_t)octIndex;
        }
    }
    /* stable sort of (prefix, internalToLeaf) by prefix (linear_octree.hpp:513-524) */
    {
        kv_t *kv = (kv_t *)malloc((size_t)nTotal * sizeof(kv_t));
        for (uint32_t i = 0; i < nTotal; i++) kv[i].key = prefixes[i], kv[i].val = internalToLeaf[i];
        stable_sort_kv(kv, nTotal);
        for (uint32_t i = 0; i < nTotal; i++) prefixes[i] = kv[i].key, internalToLeaf[i] = kv[i].val;
        free(kv);
    }
    for (uint32_t i = 0; i < nTotal; i++) {
        if (internalToLeaf[i] < 0 || (uint32_t)internalToLeaf[i] >= nTotal) goto fail;
        leafToInternal[internalToLeaf[i]] = (int32_t)i;
    }
    for (uint32_t i = 0; i < nTotal; i++) internalToLeaf[i] -= (int32_t)nInternal;

    /* getLevelRange (linear_octree.hpp:370-376) */
    size_t levelRange[MAX_DEPTH + 2];
    for (uint32_t level = 0; level <= MAX_DEPTH; level++)
        levelRange[level] = lower_bound_u64(prefixes, 0, nTotal, encode_placeholder(0, level));
    levelRange[MAX_DEPTH + 1] = nTotal;

    /* linkTree (linear_octree.hpp:379-407) */
    for (uint32_t i = 0; i < nInternal; i++) {
        size_t idxA = (size_t)leafToInternal[i];
        uint64_t prefix = prefixes[idxA];
        uint64_t nodeKey = decode_placeholder(prefix);
        uint32_t level = decode_prefix_length(prefix) / 3;
        if (level >= MAX_DEPTH) goto fail;
        uint64_t childPrefix = encode_placeholder(nodeKey, level + 1);
        size_t s = levelRange[level + 1], e = levelRange[level + 2];
        size_t childIdx = lower_bound_u64(prefixes, s, e, childPrefix);
        if (childIdx != e && childPrefix == prefixes[childIdx]) t->offsets[idxA] = (uint32_t)childIdx;
    }
    t->prefixes = prefixes;
    internal_ranges(t, internalToLeaf, 0);

    /* computeGeometry (linear_octree.hpp:557-565) */
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < nTotal; i++) {
        uint64_t prefix = prefixes[i];
        get_center(t, decode_placeholder(prefix), decode_prefix_length(prefix) / 3, &t->centers[3 * (size_t)i]);
    }
    free(internalToLeaf);
    free(leafToInternal);
    return t;
fail:
    free(prefixes);
    free(internalToLeaf);
    free(leafToInternal);
    t->prefixes = NULL;
    orc_tree_free(t);
    return NULL;
}

void orc_tree_free(orc_tree *t) {
    if (!t) return;
    free(t->leaves), free(t->counts), free(t->layout), free(t->prefixes);
    free(t->offsets), free(t->ranges), free(t->centers);
    free(t);
}
void orc_tree_info(const orc_tree *t, uint32_t info[5]) {
    info[0] = t->nLeaf, info[1] = t->nInternal, info[2] = t->nTotal, info[3] = t->maxDepthSeen, info[4] = t->rounds;
}
const uint64_t *orc_tree_leaves(const orc_tree *t) { return t->leaves; }
const uint64_t *orc_tree_counts(const orc_tree *t) { return t->counts; }
const uint64_t *orc_tree_layout(const orc_tree *t) { return t->layout; }
const uint64_t *orc_tree_prefixes(const orc_tree *t) { return t->prefixes; }
const uint32_t *orc_tree_offsets(const orc_tree *t) { return t->offsets; }
const uint64_t *orc_tree_ranges(const orc_tree *t) { return t->ranges; }
const double *orc_tree_centers(const orc_tree *t) { return t->centers; }
const double *orc_tree_radii(const orc_tree *t) { return t->radii; }

/* ------------------------------------------------------------------ kernels
 * kernel_abstract.hpp:26-45, kernel_3d.hpp, kernel_2d.hpp, util.hpp:133-137 */
enum { OUTSIDE = 0, OVERLAP = 1, INSIDE = 2 };
typedef struct {
    int type;
    double c[3];
    double radius, radiusSq;
    double bmin[3], bmax[3]; /* makeBox: center -/+ radii */
} kern_t;

static void kern_init(kern_t *k, int type, const double c[3], double radius) {
    k->type = type;
    for (int a = 0; a < 3; a++) {
        k->c[a] = c[a];
        k->bmin[a] = c[a] - radius;
        k->bmax[a] = c[a] + radius;
    }
    k->radius = radius;
    k->radiusSq = radius * radius;
}

static inline int kern_inside(const kern_t *k, const double *p) {
    switch (k->type) {
    case ORC_K_SPHERE: {
        double d1 = p[0] - k->c[0], d2 = p[1] - k->c[1], d3 = p[2] - k->c[2];
        return d1 * d1 + d2 * d2 + d3 * d3 < k->radiusSq;
    }
    case ORC_K_CIRCLE: {
        double d1 = p[0] - k->c[0], d2 = p[1] - k->c[1];
        return d1 * d1 + d2 * d2 <= k->radiusSq;
    }
    case ORC_K_CUBE:
        return p[0] >= k->bmin[0] && p[0] <= k->bmax[0] && p[1] >= k->bmin[1] && p[1] <= k->bmax[1] &&
               p[2] >= k->bmin[2] && p[2] <= k->bmax[2];
    default: /* square */
        return p[0] >= k->bmin[0] && p[0] <= k->bmax[0] && p[1] >= k->bmin[1] && p[1] <= k->bmax[1];
    }
}

static inline double dmax0(double a) { return (a < 0.0) ? 0.0 : a; } /* std::max(a, 0.0) */

/* boxIntersect(center, Vector radii): kernel_3d.hpp:181-223 (sphere), :277-302 (cube),
 * kernel_2d.hpp:107-144 (circle), :195-218 (square) */
static int kern_box_intersect(const kern_t *k, const double *nc, const double *nr) {
    switch (k->type) {
    case ORC_K_SPHERE: {
        double dx = fabs(k->c[0] - nc[0]), dy = fabs(k->c[1] - nc[1]), dz = fabs(k->c[2] - nc[2]);
        const double rx = nr[0], ry = nr[1], rz = nr[2];
        const double maxDx = k->radius + rx, maxDy = k->radius + ry, maxDz = k->radius + rz;
        if (dx > maxDx || dy > maxDy || dz > maxDz) return OUTSIDE;
        double cx = dx + rx, cy = dy + ry, cz = dz + rz;
        if ((cx * cx + cy * cy + cz * cz) <= k->radiusSq) return INSIDE;
        int numInside = (dx < rx) + (dy < ry) + (dz < rz);
        if (numInside > 1) return OVERLAP;
        dx = dmax0(dx - rx), dy = dmax0(dy - ry), dz = dmax0(dz - rz);
        if ((dx * dx + dy * dy + dz * dz) <= k->radiusSq) return OVERLAP;
        return OUTSIDE;
    }
    case ORC_K_CIRCLE: {
        double dx = fabs(k->c[0] - nc[0]), dy = fabs(k->c[1] - nc[1]);
        const double rx = nr[0], ry = nr[1];
        const double maxDx = k->radius + rx, maxDy = k->radius + ry;
        if (dx > maxDx || dy > maxDy) return OUTSIDE;
        double cx = dx + rx, cy = dy + ry;
        if ((cx * cx + cy * cy) <= k->radiusSq) return INSIDE;
        int numInside = (dx < rx) + (dy < ry);
        if (numInside > 1) return OVERLAP;
        dx = dmax0(dx - rx), dy = dmax0(dy - ry);
        if ((dx * dx + dy * dy) <= k->radiusSq) return OVERLAP;
        return OUTSIDE;
    }
    case ORC_K_CUBE: {
        const double hx = nc[0] + nr[0], lx = nc[0] - nr[0];
        const double hy = nc[1] + nr[1], ly = nc[1] - nr[1];
        const double hz = nc[2] + nr[2], lz = nc[2] - nr[2];
        if (hx < k->bmin[0] || hy < k->bmin[1] || hz < k->bmin[2] || lx > k->bmax[0] || ly > k->bmax[1] ||
            lz > k->bmax[2])
            return OUTSIDE;
        if (hx <= k->bmax[0] && hy <= k->bmax[1] && hz <= k->bmax[2] && lx >= k->bmin[0] && ly >= k->bmin[1] &&
            lz >= k->bmin[2])
            return INSIDE;
        return OVERLAP;
    }
    default: {
        const double hx = nc[0] + nr[0], lx = nc[0] - nr[0];
        const double hy = nc[1] + nr[1], ly = nc[1] - nr[1];
        if (hx < k->bmin[0] || hy < k->bmin[1] || lx > k->bmax[0] || ly > k->bmax[1]) return OUTSIDE;
        if (hx <= k->bmax[0] && hy <= k->bmax[1] && lx >= k->bmin[0] && ly >= k->bmin[1]) return INSIDE;
        return OVERLAP;
    }
    }
}

/* boxOverlap(center, Vector radii): Kernel3D kernel_3d.hpp:100-115, Kernel2D kernel_2d.hpp:29-41 */
static int kern_box_overlap(const kern_t *k, const double *nc, const double *nr) {
    int is3d = (k->type == ORC_K_SPHERE || k->type == ORC_K_CUBE);
    if (nc[0] + nr[0] < k->bmin[0] || nc[1] + nr[1] < k->bmin[1] || (is3d && nc[2] + nr[2] < k->bmin[2]))
        return 0;
    if (nc[0] - nr[0] > k->bmax[0] || nc[1] - nr[1] > k->bmax[1] || (is3d && nc[2] - nr[2] > k->bmax[2]))
        return 0;
    return 1;
}

/* ------------------------------------------------------------------ pointer octree (competitor)
 * The reference's pointer-based Octree (inc/structures/octree.hpp:20-396, src/octree.cpp:17-322),
 * restated as a CPU baseline: built by inserting the points one by one (buildOctree -> insertPoints,
 * octree.cpp:207-243); a leaf already holding maxPointsLeaf points splits into 8 octants
 * (createOctants, :245-257: child i has x +/- radii.x * 0.5F by bit 4, y by bit 2, z by bit 1, radii
 * halved) and hands its points to them (fillOctants, :259-270; octantIdx, :272-283: x >= centre.x ->
 * bit 4, ...). Searches (neighbors, octree.hpp:156-188): depth first over the octants whose box
 * passes the kernel's boxOverlap, every point of a reached leaf tested with isInside. The reference
 * splits with no depth limit (MIN_OCTANT_RADIUS = 0), so more than maxPointsLeaf identical points
 * would recurse without end; here splitting stops at depth 64 (the leaf then keeps every point). */
typedef struct {
    double c[3], r[3];
    int64_t child;  /* index of the first of 8 children, -1 for a leaf */
    uint32_t *pts, n, cap;
} pnode_t;

struct orc_ptree {
    pnode_t *nodes;
    size_t nn, capn;
    const double *xyz;
    uint32_t max_leaf;
    uint32_t depth;
};

static size_t ptree_new_node(orc_ptree *t, const double c[3], const double r[3]) {
    if (t->nn == t->capn) {
        t->capn = t->capn ? 2 * t->capn : 1024;
        t->nodes = (pnode_t *)realloc(t->nodes, t->capn * sizeof(pnode_t));
    }
    pnode_t *p = &t->nodes[t->nn];
    memcpy(p->c, c, sizeof p->c);
    memcpy(p->r, r, sizeof p->r);
    p->child = -1, p->pts = NULL, p->n = p->cap = 0;
    return t->nn++;
}
static void ptree_push(pnode_t *p, uint32_t i) {
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 8;
        p->pts = (uint32_t *)realloc(p->pts, p->cap * sizeof(uint32_t));
    }
    p->pts[p->n++] = i;
}
static inline unsigned ptree_octant(const orc_ptree *t, size_t node, uint32_t i) {
    const double *p = &t->xyz[3 * (size_t)i];
    const pnode_t *n = &t->nodes[node];
    return (p[0] >= n->c[0] ? 4u : 0u) | (p[1] >= n->c[1] ? 2u : 0u) | (p[2] >= n->c[2] ? 1u : 0u);
}
static void ptree_insert(orc_ptree *t, size_t node, uint32_t i, uint32_t depth) {
    for (;;) {
        pnode_t *n = &t->nodes[node];
        if (n->child >= 0) {
            node = (size_t)n->child + ptree_octant(t, node, i);
            depth++;
            continue;
        }
        if (n->n == 0 || n->n < t->max_leaf || depth >= 64) {
            ptree_push(n, i);
            if (depth > t->depth) t->depth = depth;
            return;
        }
        /* createOctants + fillOctants, then insert into the octant of i */
        double c[3] = {n->c[0], n->c[1], n->c[2]}, r[3] = {n->r[0], n->r[1], n->r[2]};
        const size_t first = t->nn;
        for (unsigned o = 0; o < 8; o++) {
            double cc[3], rr[3];
            cc[0] = c[0] + r[0] * ((o & 4u) ? 0.5f : -0.5f);
            cc[1] = c[1] + r[1] * ((o & 2u) ? 0.5f : -0.5f);
            cc[2] = c[2] + r[2] * ((o & 1u) ? 0.5f : -0.5f);
            rr[0] = r[0] * 0.5f, rr[1] = r[1] * 0.5f, rr[2] = r[2] * 0.5f;
            ptree_new_node(t, cc, rr);
        }
        n = &t->nodes[node]; /* the array may have moved */
        n->child = (int64_t)first;
        uint32_t *old = n->pts, cnt = n->n;
        n->pts = NULL, n->n = n->cap = 0;
        for (uint32_t k = 0; k < cnt; k++) ptree_insert(t, first + ptree_octant(t, node, old[k]), old[k], depth + 1);
        free(old);
        node = first + ptree_octant(t, node, i);
        depth++;
    }
}

orc_ptree *orc_ptree_build(const double *xyz, size_t n, uint32_t max_points_leaf, const double center[3],
                           const double radii[3]) {
    orc_ptree *t = (orc_ptree *)calloc(1, sizeof(orc_ptree));
    t->xyz = xyz;
    t->max_leaf = max_points_leaf;
    ptree_new_node(t, center, radii);
    for (size_t i = 0; i < n; i++) ptree_insert(t, 0, (uint32_t)i, 0);
    return t;
}

void orc_ptree_free(orc_ptree *t) {
    if (!t) return;
    for (size_t i = 0; i < t->nn; i++) free(t->nodes[i].pts);
    free(t->nodes);
    free(t);
}

void orc_ptree_info(const orc_ptree *t, uint64_t info[3]) {
    uint64_t leaves = 0;
    for (size_t i = 0; i < t->nn; i++) leaves += t->nodes[i].child < 0;
    info[0] = leaves, info[1] = t->nn - leaves, info[2] = t->depth;
}

/* neighbors(kernel, condition) (octree.hpp:156-188) for nq centres of one radius; sizes out, and the
 * index lists when out_idx != NULL (offsets from the exclusive prefix of sizes, filled by the caller's
 * second call). Returns the total. */
uint64_t orc_ptree_radius(const orc_ptree *t, int kernel, const double *centers, size_t nq, double radius,
                          uint64_t *sizes, const uint64_t *offsets, uint64_t *out_idx) {
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        size_t cap = 256, sp;
        size_t *stack = (size_t *)malloc(cap * sizeof(size_t));
#pragma omp for schedule(dynamic, 64)
        for (size_t q = 0; q < nq; q++) {
            kern_t k;
            kern_init(&k, kernel, &centers[3 * q], radius);
            uint64_t cnt = 0;
            sp = 0;
            stack[sp++] = 0;
            while (sp) {
                const pnode_t *n = &t->nodes[stack[--sp]];
                if (n->child < 0) {
                    for (uint32_t j = 0; j < n->n; j++)
                        if (kern_inside(&k, &t->xyz[3 * (size_t)n->pts[j]])) {
                            if (out_idx) out_idx[offsets[q] + cnt] = n->pts[j];
                            cnt++;
                        }
                } else {
                    if (sp + 8 > cap) cap *= 2, stack = (size_t *)realloc(stack, cap * sizeof(size_t));
                    for (unsigned o = 0; o < 8; o++) {
                        const pnode_t *ch = &t->nodes[n->child + o];
                        if (kern_box_overlap(&k, ch->c, ch->r)) stack[sp++] = (size_t)n->child + o;
                    }
                }
            }
            sizes[q] = cnt;
            total += cnt;
        }
        free(stack);
    }
    return total;
}

/* ------------------------------------------------------------------ traversal + searches */
typedef struct {
    uint64_t *v;
    size_t n, cap;
} vec_t;
static void vpush(vec_t *v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->v = (uint64_t *)realloc(v->v, v->cap * sizeof(uint64_t));
    }
    v->v[v->n++] = x;
}

typedef struct {
    const orc_tree *t;
    const double *xyz;
    kern_t k;
    int algo;
    uint32_t level;  /* neighborsApprox: precision level and bound mode */
    int upper;
    int count_only;
    vec_t out;
    uint64_t size;
    uint64_t tested, classified;
} search_t;

static void add_range(search_t *s, uint64_t first, uint64_t last) { /* NeighborSet::addRange */
    if (first <= last) {
        if (!s->count_only) vpush(&s->out, first), vpush(&s->out, last);
        s->size += last - first;
    }
}

/* continuationCriterion of the variants (linear_octree.hpp:698-714, 749-769, 800-802; neighborsApprox
 * 984-1005: beyond the precision level a node is added whole in upper-bound mode and pruned either way) */
static int criterion(search_t *s, uint32_t node, uint32_t depth) {
    const orc_tree *t = s->t;
    s->classified++;
    if (s->algo == ORC_A_NEIGHBORS) return kern_box_overlap(&s->k, &t->centers[3 * (size_t)node], &t->radii[3 * depth]);
    if (s->algo == ORC_A_APPROX && depth > s->level) {
        if (s->upper) add_range(s, t->ranges[2 * (size_t)node], t->ranges[2 * (size_t)node + 1]);
        return 0;
    }
    int j = kern_box_intersect(&s->k, &t->centers[3 * (size_t)node], &t->radii[3 * depth]);
    if (j == INSIDE) {
        uint64_t a = t->ranges[2 * (size_t)node], b = t->ranges[2 * (size_t)node + 1];
        if (s->algo == ORC_A_STRUCT || s->algo == ORC_A_APPROX) {
            add_range(s, a, b);
        } else {
            for (uint64_t i = a; i < b; ++i) {
                if (!s->count_only) vpush(&s->out, i);
                s->size++;
            }
        }
        return 0;
    }
    return j == OVERLAP;
}

/* endpointAction of the three variants (linear_octree.hpp:716-732, 771-780, 804-813) */
static void endpoint(search_t *s, uint32_t node) {
    const orc_tree *t = s->t;
    uint64_t a = t->ranges[2 * (size_t)node], b = t->ranges[2 * (size_t)node + 1];
    s->tested += b - a;
    if (s->algo == ORC_A_STRUCT || s->algo == ORC_A_APPROX) {
        uint64_t rs = a;
        for (uint64_t i = a; i < b; ++i) {
            if (!kern_inside(&s->k, &s->xyz[3 * i])) {
                if (rs < i) add_range(s, rs, i);
                rs = i + 1;
            }
        }
        if (rs < b) add_range(s, rs, b);
    } else {
        for (uint64_t i = a; i < b; ++i)
            if (kern_inside(&s->k, &s->xyz[3 * i])) {
                if (!s->count_only) vpush(&s->out, i);
                s->size++;
            }
    }
}

/* singleTraversal (linear_octree.hpp:653-686). The stack is sized for the worst case (the
 * reference's 128-entry stack overflows silently beyond that). */
static void single_traversal(search_t *s) {
    const orc_tree *t = s->t;
    if (!criterion(s, 0, 0)) return;
    if (t->offsets[0] == 0) {
        endpoint(s, 0);
        return;
    }
    uint32_t stackN[8 * 23 + 8], stackD[8 * 23 + 8];
    stackN[0] = 0, stackD[0] = 0;
    uint32_t sp = 1, node = 0, depth = 0;
    do {
        for (uint32_t oct = 0; oct < 8; ++oct) {
            uint32_t child = t->offsets[node] + oct;
            if (criterion(s, child, depth + 1)) {
                if (t->offsets[child] == 0)
                    endpoint(s, child);
                else
                    stackN[sp] = child, stackD[sp] = depth + 1, sp++;
            }
        }
        --sp;
        node = stackN[sp], depth = stackD[sp];
    } while (node != 0);
}

struct orc_result {
    size_t nq;
    uint64_t *offsets, *entries, *sizes;
};

/* Cube / square kernel from a Vector of half extents: KernelAbstract(center, radii) + makeBox
 * (kernel_abstract.hpp:39-45), box = centre -/+ radii per axis. */
static void kern_init3(kern_t *k, int type, const double c[3], const double r3[3]) {
    kern_init(k, type, c, r3[0]);
    for (int a = 0; a < 3; a++) {
        k->bmin[a] = c[a] - r3[a];
        k->bmax[a] = c[a] + r3[a];
    }
}

static orc_result *radius_any(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                              const double *radii, const double *radii3, size_t nq, uint64_t *stats, uint32_t level,
                              int upper);

orc_result *orc_radius(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                       const double *radii, size_t nq, uint64_t *stats) {
    return radius_any(t, xyz, kernel, algo, centers, radii, NULL, nq, stats, 0, 0);
}

orc_result *orc_radius_box(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                           const double *radii3, size_t nq) {
    return radius_any(t, xyz, kernel, algo, centers, NULL, radii3, nq, NULL, 0, 0);
}

orc_result *orc_radius_approx(const orc_tree *t, const double *xyz, int kernel, const double *centers,
                              const double *radii, const double *radii3, size_t nq, uint32_t level, int upper) {
    return radius_any(t, xyz, kernel, ORC_A_APPROX, centers, radii, radii3, nq, NULL, level, upper);
}

static orc_result *radius_any(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                              const double *radii, const double *radii3, size_t nq, uint64_t *stats, uint32_t level,
                              int upper) {
    vec_t *outs = (vec_t *)calloc(nq ? nq : 1, sizeof(vec_t));
    orc_result *r = (orc_result *)calloc(1, sizeof(orc_result));
    r->nq = nq;
    r->sizes = (uint64_t *)calloc(nq ? nq : 1, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < nq; q++) {
        search_t s;
        memset(&s, 0, sizeof(s));
        s.t = t, s.xyz = xyz, s.algo = algo, s.level = level, s.upper = upper;
        if (radii3)
            kern_init3(&s.k, kernel, &centers[3 * q], &radii3[3 * q]);
        else
            kern_init(&s.k, kernel, &centers[3 * q], radii[q]);
        single_traversal(&s);
        outs[q] = s.out;
        r->sizes[q] = s.size;
        if (stats) stats[2 * q] = s.tested, stats[2 * q + 1] = s.classified;
    }
    r->offsets = (uint64_t *)malloc((nq + 1) * sizeof(uint64_t));
    r->offsets[0] = 0;
    for (size_t q = 0; q < nq; q++) r->offsets[q + 1] = r->offsets[q] + outs[q].n;
    r->entries = (uint64_t *)malloc((r->offsets[nq] ? r->offsets[nq] : 1) * sizeof(uint64_t));
    for (size_t q = 0; q < nq; q++) {
        if (outs[q].n) memcpy(r->entries + r->offsets[q], outs[q].v, outs[q].n * sizeof(uint64_t));
        free(outs[q].v);
    }
    free(outs);
    return r;
}

uint64_t orc_radius_count(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                          const double *radii, size_t nq, uint64_t *sizes) {
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (size_t q = 0; q < nq; q++) {
        search_t s;
        memset(&s, 0, sizeof(s));
        s.t = t, s.xyz = xyz, s.algo = algo, s.count_only = 1;
        kern_init(&s.k, kernel, &centers[3 * q], radii[q]);
        single_traversal(&s);
        if (sizes) sizes[q] = s.size;
        total += s.size;
    }
    return total;
}

/* What the reference harness does per query (neighbor_benchmarks.hpp:533-546): build the result
 * (ranges for Struct, index vector otherwise), keep only its size, drop it. Timed CPU baseline. */
uint64_t orc_radius_harness(const orc_tree *t, const double *xyz, int kernel, int algo, const double *centers,
                            const double *radii, size_t nq) {
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (size_t q = 0; q < nq; q++) {
        search_t s;
        memset(&s, 0, sizeof(s));
        s.t = t, s.xyz = xyz, s.algo = algo;
        kern_init(&s.k, kernel, &centers[3 * q], radii[q]);
        single_traversal(&s);
        total += s.size;
        free(s.out.v);
    }
    return total;
}

size_t orc_result_total(const orc_result *r) { return r->offsets[r->nq]; }
const uint64_t *orc_result_offsets(const orc_result *r) { return r->offsets; }
const uint64_t *orc_result_entries(const orc_result *r) { return r->entries; }
const uint64_t *orc_result_sizes(const orc_result *r) { return r->sizes; }
void orc_result_free(orc_result *r) {
    if (!r) return;
    free(r->offsets), free(r->entries), free(r->sizes), free(r);
}

orc_result *orc_radius_brute(const double *xyz, size_t n, int kernel, const double *centers, const double *radii,
                             size_t nq) {
    vec_t *outs = (vec_t *)calloc(nq ? nq : 1, sizeof(vec_t));
    orc_result *r = (orc_result *)calloc(1, sizeof(orc_result));
    r->nq = nq;
    r->sizes = (uint64_t *)calloc(nq ? nq : 1, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < nq; q++) {
        kern_t k;
        kern_init(&k, kernel, &centers[3 * q], radii[q]);
        for (size_t i = 0; i < n; i++)
            if (kern_inside(&k, &xyz[3 * i])) vpush(&outs[q], i);
        r->sizes[q] = outs[q].n;
    }
    r->offsets = (uint64_t *)malloc((nq + 1) * sizeof(uint64_t));
    r->offsets[0] = 0;
    for (size_t q = 0; q < nq; q++) r->offsets[q + 1] = r->offsets[q] + outs[q].n;
    r->entries = (uint64_t *)malloc((r->offsets[nq] ? r->offsets[nq] : 1) * sizeof(uint64_t));
    for (size_t q = 0; q < nq; q++) {
        if (outs[q].n) memcpy(r->entries + r->offsets[q], outs[q].v, outs[q].n * sizeof(uint64_t));
        free(outs[q].v);
    }
    free(outs);
    return r;
}

/* ------------------------------------------------------------------ kNN */
typedef struct {
    double d;
    uint64_t index; /* point index or node index */
    uint32_t depth;
    uint32_t isOctant;
} hent_t;
typedef struct {
    hent_t *a;
    size_t n, cap;
} heap_t;
static void hpush(heap_t *h, hent_t e) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 512;
        h->a = (hent_t *)realloc(h->a, h->cap * sizeof(hent_t));
    }
    size_t i = h->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!(e.d < h->a[p].d)) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i] = e;
}
static hent_t hpop(heap_t *h) {
    hent_t top = h->a[0], last = h->a[--h->n];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        double md = last.d;
        if (l < h->n && h->a[l].d < md) m = l, md = h->a[l].d;
        if (r < h->n && h->a[r].d < md) m = r;
        if (m == i) break;
        h->a[i] = h->a[m];
        i = m;
    }
    if (h->n) h->a[i] = last;
    return top;
}

/* distPointsSquared (linear_octree.hpp:844-851) */
static inline double dist_pp(const double *p, const double *q) {
    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}
/* distPointOctantSquared (linear_octree.hpp:854-878) */
static inline double dist_po(const double *p, const double *c, const double *r) {
    double mn[3], mx[3], cl[3];
    for (int a = 0; a < 3; a++) mn[a] = c[a] - r[a], mx[a] = c[a] + r[a];
    for (int a = 0; a < 3; a++) {
        double v = (mx[a] < p[a]) ? mx[a] : p[a]; /* std::min(p, max) */
        cl[a] = (mn[a] < v) ? v : mn[a];          /* std::max(min, v) */
    }
    double dx = p[0] - cl[0], dy = p[1] - cl[1], dz = p[2] - cl[2];
    return dx * dx + dy * dy + dz * dz;
}

/* knn (linear_octree.hpp:899-942): best-first over a min-heap of octants and points; stops after
 * min(n, k) points are popped. Ties between equal distances pop in heap order (unspecified). */
uint64_t orc_knn(const orc_tree *t, const double *xyz, size_t n, const double *centers, size_t nq, uint32_t k,
                 int64_t *idx, double *d2) {
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : total)
    for (size_t q = 0; q < nq; q++) {
        const double *p = &centers[3 * q];
        heap_t h = {0, 0, 0};
        hent_t root = {0.0, 0, 0, 1};
        hpush(&h, root);
        size_t inserted = 0, maxPoints = n < k ? n : k;
        for (uint32_t j = 0; j < k; j++) idx[q * k + j] = -1, d2[q * k + j] = NAN;
        while (inserted < maxPoints) {
            hent_t top = hpop(&h);
            if (top.isOctant) {
                uint32_t node = (uint32_t)top.index;
                if (t->offsets[node] == 0) {
                    for (uint64_t i = t->ranges[2 * (size_t)node]; i < t->ranges[2 * (size_t)node + 1]; ++i) {
                        hent_t e = {dist_pp(p, &xyz[3 * i]), i, 0, 0};
                        hpush(&h, e);
                    }
                } else {
                    uint32_t cd = top.depth + 1;
                    for (uint32_t oct = 0; oct < 8; ++oct) {
                        uint32_t c = t->offsets[node] + oct;
                        hent_t e = {dist_po(p, &t->centers[3 * (size_t)c], &t->radii[3 * cd]), c, cd, 1};
                        hpush(&h, e);
                    }
                }
            } else {
                idx[q * k + inserted] = (int64_t)top.index;
                d2[q * k + inserted] = top.d;
                inserted++;
            }
        }
        free(h.a);
        total += inserted;
    }
    return total;
}

typedef struct {
    double d;
    uint64_t i;
} di_t;
static int di_cmp(const void *a, const void *b) {
    const di_t *x = (const di_t *)a, *y = (const di_t *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

uint64_t orc_knn_brute(const double *xyz, size_t n, const double *centers, size_t nq, uint32_t k, int64_t *idx,
                       double *d2) {
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        di_t *buf = (di_t *)malloc((n ? n : 1) * sizeof(di_t));
#pragma omp for schedule(dynamic, 4)
        for (size_t q = 0; q < nq; q++) {
            for (size_t i = 0; i < n; i++) buf[i].d = dist_pp(&centers[3 * q], &xyz[3 * i]), buf[i].i = i;
            qsort(buf, n, sizeof(di_t), di_cmp);
            size_t m = n < k ? n : k;
            for (uint32_t j = 0; j < k; j++) {
                idx[q * k + j] = j < m ? (int64_t)buf[j].i : -1;
                d2[q * k + j] = j < m ? buf[j].d : NAN;
            }
            total += m;
        }
        free(buf);
    }
    return total;
}

/* ------------------------------------------------------------------ LAS ingest */

#include <stdio.h>

static uint64_t las_u(const unsigned char *p, int bytes) { /* little-endian unsigned field */
    uint64_t v = 0;
    for (int i = bytes - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

static double las_f64(const unsigned char *p) {
    uint64_t bits = las_u(p, 8);
    double d;
    memcpy(&d, &bits, 8);
    return d;
}

/* Public header block, LAS 1.4 R15 table 3 (1.2 and 1.3 are its first 227 and 235 bytes). LASlib's
 * header.number_of_point_records is the legacy 32-bit count (offset 107); for a LAS 1.4 file whose
 * legacy count is 0 the 64-bit count (offset 247) is used. */
int orc_las_header(const char *path, uint64_t info[5], double so[6]) {
    unsigned char h[375];
    memset(h, 0, sizeof(h));
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    size_t got = fread(h, 1, sizeof(h), f);
    fclose(f);
    if (got < 227 || memcmp(h, "LASF", 4) != 0) return -1;
    unsigned fmt = h[104];
    if (fmt > 10) return -2; /* LAZ sets bits 6-7 */
    uint64_t n = las_u(h + 107, 4);
    if (h[25] >= 4 && n == 0 && got >= 255) n = las_u(h + 247, 8);
    info[0] = h[25];
    info[1] = fmt;
    info[2] = las_u(h + 105, 2);
    info[3] = n;
    info[4] = las_u(h + 96, 4);
    for (int a = 0; a < 3; ++a) {
        so[a] = las_f64(h + 131 + 8 * a);
        so[3 + a] = las_f64(h + 155 + 8 * a);
    }
    return 0;
}

/* Point records, LAS 1.4 R15 tables 7-17. Formats 0-5: intensity @12, return bits @14 (return number
 * 0-2, number of returns 3-5, scan direction 6, edge 7), classification @15 (LASlib's 5-bit class),
 * scan angle rank @16 (int8), user data @17, point source id @18, RGB @20 (format 2) or @28 (3, 5).
 * Formats 6-10: intensity @12, return number (bits 0-3) and number of returns (4-7) @14, scan
 * direction (bit 6) and edge (bit 7) @15, classification @16, user data @17, scan angle @18 (int16,
 * 0.006 degree units), point source id @20, RGB @30 (7, 8, 10); LASlib fills its legacy fields from
 * them (returns above 7 folded: number of returns 7, return number 7 when it is the last else 6;
 * classification 0 when above 31), and the reference casts get_scan_angle() = 0.006f * angle to char. */
int orc_las_read(const char *path, double *xyz, double *intensity, uint8_t *fields, uint16_t *psid, uint32_t *rgb) {
    uint64_t info[5];
    double so[6];
    int r = orc_las_header(path, info, so);
    if (r) return r;
    const unsigned fmt = (unsigned)info[1], rl = (unsigned)info[2];
    const uint64_t n = info[3];
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    if (fseek(f, (long)info[4], SEEK_SET) != 0) {
        fclose(f);
        return -1;
    }
    unsigned char *rec = (unsigned char *)malloc(rl ? rl : 1);
    for (uint64_t i = 0; i < n; ++i) {
        if (fread(rec, 1, rl, f) != rl) {
            free(rec);
            fclose(f);
            return -1;
        }
        const int32_t X = (int32_t)(uint32_t)las_u(rec, 4), Y = (int32_t)(uint32_t)las_u(rec + 4, 4);
        const int32_t Z = (int32_t)(uint32_t)las_u(rec + 8, 4);
        xyz[3 * i + 0] = X * so[0] + so[3];
        xyz[3 * i + 1] = Y * so[1] + so[4];
        xyz[3 * i + 2] = Z * so[2] + so[5];
        intensity[i] = (double)las_u(rec + 12, 2);
        uint8_t *fl = fields + 8 * i;
        unsigned rgb_at = 0;
        if (fmt < 6) {
            const unsigned b = rec[14];
            fl[0] = b & 7u;
            fl[1] = (b >> 3) & 7u;
            fl[2] = (b >> 6) & 1u;
            fl[3] = (b >> 7) & 1u;
            fl[4] = rec[15] & 31u;
            fl[5] = rec[16];
            fl[6] = rec[17];
            psid[i] = (uint16_t)las_u(rec + 18, 2);
            if (fmt == 2) rgb_at = 20;
            if (fmt == 3 || fmt == 5) rgb_at = 28;
        } else {
            const unsigned rn = rec[14] & 15u, nr = rec[14] >> 4;
            if (nr <= 7) {
                fl[0] = (uint8_t)(rn & 7u); /* LASlib's 3-bit field keeps the low bits */
                fl[1] = (uint8_t)nr;
            } else {
                fl[0] = (uint8_t)(rn <= 6 ? rn : (rn >= nr ? 7 : 6));
                fl[1] = 7;
            }
            fl[2] = (rec[15] >> 6) & 1u;
            fl[3] = (rec[15] >> 7) & 1u;
            fl[4] = rec[16] <= 31 ? rec[16] : 0;
            const int16_t ang = (int16_t)(uint16_t)las_u(rec + 18, 2);
            const float deg = 0.006f * (float)ang;
            fl[5] = (uint8_t)(int8_t)(int)deg;
            fl[6] = rec[17];
            psid[i] = (uint16_t)las_u(rec + 20, 2);
            if (fmt == 7 || fmt == 8 || fmt == 10) rgb_at = 30;
        }
        fl[7] = 0;
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = rgb_at ? (uint32_t)las_u(rec + rgb_at + 2 * c, 2) : 0u;
    }
    free(rec);
    fclose(f);
    return 0;
}
