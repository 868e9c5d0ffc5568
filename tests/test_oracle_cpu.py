"""Pins the CPU oracle (oracle/oracle.c) without a GPU: against the reference's own libmorton
(golden vectors), the reference's known-answer tests, the platform's own double -> uint64
conversion, and what the mathematics fixes (bijections, curve adjacency, closed-form tree
structure, brute-force neighbour sets and kNN). Each check is chosen so that a plausible mistake in
the restatement (a dropped term, a wrong index, a transposed operand, a wrong strictness) fails it."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MAXC = (1 << 21) - 1
UPPER = 1 << 63


# ------------------------------------------------------------------ a2 quantisation
X86_CVT_SRC = r"""
#include <stdint.h>
#include <stdio.h>
#include <string.h>
/* The reference's conversion (point_encoder.hpp:46-48): std::min((coords_t)(x_t * 2^21), maxCoord),
 * coords_t = uint_fast32_t (64-bit on x86-64 glibc). Values arrive as bit patterns at run time so
 * nothing is constant-folded. */
int main(void) {
    unsigned long long bits;
    while (scanf("%llx", &bits) == 1) {
        double t;
        memcpy(&t, &bits, 8);
        uint_fast32_t q = (uint_fast32_t)(t * (1 << 21));
        uint_fast32_t m = (1u << 21) - 1u;
        printf("%llu\n", (unsigned long long)(q < m ? q : m));
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_quantize_matches_platform_conversion():
    """getAnchorCoords (point_encoder.hpp:37-49) against the same conversion compiled by gcc with the
    reference's Release flags (cmake/CMakeConfig.cmake:41,44): the transform
    x_t = ((p - c) + r) / (2r) is evaluated in numpy (IEEE FP64, same operations), the conversion to
    the 64-bit coords_t and the clamp by the compiled program, on boundary and special inputs."""
    nan, inf = float("nan"), float("inf")
    cases = [(0.0, 5.0, 5.0), (10.0, 5.0, 5.0), (5.0, 5.0, 5.0), (1.0, 1.5, 1.5), (-1e-9, 5.0, 5.0),
             (-1e-300, 5.0, 5.0), (-5.0, 5.0, 5.0), (-10.0, 5.0, 5.0), (-1e6, 5.0, 5.0), (20.0, 5.0, 5.0),
             (1e15, 5.0, 5.0), (1e300, 5.0, 5.0), (2.0, 2.0, 0.0), (3.0, 2.0, 0.0), (1.0, 2.0, 0.0),
             (1e300, 0.0, 1e-300), (-1e300, 0.0, 1e-300), (nan, 1.0, 1.0), (inf, 1.0, 1.0), (-inf, 1.0, 1.0),
             (1.0, nan, 1.0), (10.0 - 1e-15, 5.0, 5.0), (3.0, -2.0, 2.5), (-2.0 - 2.5, -2.0, 2.5)]
    rng = np.random.default_rng(5)
    for _ in range(300):
        c, r = rng.uniform(-1e6, 1e6), abs(rng.normal()) * 10.0 ** rng.integers(-3, 6)
        cases.append((c + r * rng.uniform(-1.05, 1.05), c, r))
    with np.errstate(all="ignore"):
        ts = [((np.float64(p) - np.float64(c)) + np.float64(r)) / (2 * np.float64(r)) for p, c, r in cases]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "cvt.c"), os.path.join(d, "cvt")
        open(src, "w").write(X86_CVT_SRC)
        subprocess.run(["gcc", "-O2", "-mavx2", "-mbmi2", "-o", exe, src], check=True)
        inp = "\n".join(f"{np.float64(t).view(np.uint64):x}" for t in ts) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    for (p, c, r), t, w in zip(cases, ts, out):
        assert O.lib().orc_quantize_axis(p, c, r) == int(w), f"p={p!r} c={c!r} r={r!r} t={t!r}"


def test_quantize_closed_form_edges():
    """p = min' -> 0, p = max' -> 2^21 - 1 (clamped, point_encoder.hpp:45), interior truncation."""
    q = O.lib().orc_quantize_axis
    assert q(0.0, 5.0, 5.0) == 0
    assert q(10.0, 5.0, 5.0) == MAXC
    assert q(5.0, 5.0, 5.0) == 1 << 20
    # 1/3 of the way: floor(2^21 / 3)
    assert q(1.0, 1.5, 1.5) == (1 << 21) // 3
    # degenerate axis (r = 0): 0/0 is NaN, which the x86 conversion sends to 2^63 -> clamp
    assert q(2.0, 2.0, 0.0) == MAXC


# ------------------------------------------------------------------ a3 Morton
def test_morton_matches_reference_libmorton_golden():
    """Oracle Morton keys against keys computed by the reference's vendored libmorton
    (tests/golden/make_morton_golden.py, from oracle/_ref/libref_morton.so)."""
    g = np.load(os.path.join(HERE, "golden", "morton_ref.npz"))
    for (x, y, z), k in zip(g["xyz"], g["keys"]):
        assert O.morton_encode(x, y, z) == int(k)
        assert O.morton_decode(k) == (int(x), int(y), int(z))


def test_morton_live_reference_when_built():
    so = os.path.join(ROOT, "oracle", "_ref", "libref_morton.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref not built here")
    import ctypes as C
    L = C.CDLL(so)
    L.ref_morton3d_encode.restype = C.c_uint64
    L.ref_morton3d_encode.argtypes = [C.c_uint64] * 3
    rng = np.random.default_rng(77)
    for x, y, z in rng.integers(0, 1 << 21, size=(20000, 3)):
        assert O.morton_encode(x, y, z) == L.ref_morton3d_encode(int(x), int(y), int(z))


def test_morton_kat_encoder_decoder_consistency():  # test_complex.cpp:104-122
    c1 = O.morton_encode(1000, 2000, 3000)
    assert O.morton_decode(c1) == (1000, 2000, 3000)
    assert O.morton_encode(*O.morton_decode(c1)) == c1


def test_morton_bit_placement():
    """x bit j -> key bit 3j, y -> 3j+1, z -> 3j+2 (a transposed axis fails here)."""
    for j in range(21):
        assert O.morton_encode(1 << j, 0, 0) == 1 << (3 * j)
        assert O.morton_encode(0, 1 << j, 0) == 1 << (3 * j + 1)
        assert O.morton_encode(0, 0, 1 << j) == 1 << (3 * j + 2)


# ------------------------------------------------------------------ a4 Hilbert
def test_hilbert_encode_inverts_reference_decode():
    """The oracle's encode (hilbert_encoder_3d.hpp:29-66) and decode (:256-289) restate two separate
    blocks of the reference; each must invert the other. A convention slip in either one (a wrong
    complement mask or rotation) breaks this."""
    rng = np.random.default_rng(3)
    pts = list(rng.integers(0, 1 << 21, size=(5000, 3)))
    pts += [(0, 0, 0), (MAXC, MAXC, MAXC), (MAXC, 0, 0), (0, MAXC, 0), (0, 0, MAXC)]
    pts += [(1 << b, 1 << ((b + 7) % 21), 1 << ((b + 13) % 21)) for b in range(21)]
    for x, y, z in pts:
        k = O.hilbert_encode(x, y, z)
        assert k < UPPER
        assert O.hilbert_decode(k) == (int(x), int(y), int(z))
    for k in rng.integers(0, 1 << 63, size=5000, dtype=np.uint64):
        assert O.hilbert_encode(*O.hilbert_decode(int(k))) == int(k)


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_hilbert_coarse_curve_is_a_hilbert_curve(level):
    """The level-L prefixes of the 8^L cells form a bijection onto [0, 8^L), the first cell is the
    origin, and consecutive cells are face neighbours (unit Manhattan steps)."""
    s = 21 - level
    n = 1 << level
    order = np.full(8 ** level, -1, np.int64)
    cells = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                pref = O.hilbert_encode(x << s, y << s, z << s) >> (3 * s)
                assert order[pref] == -1
                order[pref] = 1
                cells[pref] = (x, y, z)
    assert (order == 1).all()
    assert cells[0] == (0, 0, 0)
    for k in range(8 ** level - 1):
        a, b = np.array(cells[k]), np.array(cells[k + 1])
        assert np.abs(a - b).sum() == 1, (level, k, cells[k], cells[k + 1])


def test_hilbert_level1_digit_table():
    """First level: octant (xi, yi, zi) gets digit m2h[(xi<<2)|(yi<<1)|zi] with
    m2h = {0,1,3,2,7,6,4,5} (hilbert_encoder_3d.hpp:22)."""
    m2h = [0, 1, 3, 2, 7, 6, 4, 5]
    for o in range(8):
        x, y, z = (o >> 2) & 1, (o >> 1) & 1, o & 1
        assert O.hilbert_encode(x << 20, y << 20, z << 20) >> 60 == m2h[o]


def test_hilbert_full_depth_adjacency():
    rng = np.random.default_rng(9)
    for k in rng.integers(0, (1 << 63) - 1, size=3000, dtype=np.uint64):
        a, b = np.array(O.hilbert_decode(int(k))), np.array(O.hilbert_decode(int(k) + 1))
        assert np.abs(a - b).sum() == 1


# ------------------------------------------------------------------ a1 bbox + reference KATs
def test_mbb_kat():  # test_geometry.cpp:27-40
    c, r = O.mbb([[0, 0, 0], [10, 10, 10], [5, 5, 5]])
    assert tuple(c) == (5.0, 5.0, 5.0) and r[0] == 5.0


def test_mbb_closed_form():
    rng = np.random.default_rng(1)
    xyz = rng.normal(size=(1000, 3)) * [3, 50, 0.1] + [1e5, -7, 3]
    c, r = O.mbb(xyz)
    mn, mx = xyz.min(0), xyz.max(0)
    np.testing.assert_array_equal(c, mn + (mx - mn) / 2)
    np.testing.assert_array_equal(r, (mx - mn) / 2)


PTS4 = [[1.0, 1.0, 1.0], [9.0, 9.0, 9.0], [5.0, 5.0, 5.0], [1.0, 9.0, 1.0]]


@pytest.mark.parametrize("enc", [O.ENC_MORTON, O.ENC_HILBERT])
def test_sorting_kat(enc):  # test_encoders.cpp:29-65
    sx, keys, perm, c, r = O.sort_points(enc, PTS4)
    assert len(keys) == 4
    np.testing.assert_array_equal(c - r, [1, 1, 1])
    np.testing.assert_array_equal(c + r, [9, 9, 9])
    if enc == O.ENC_MORTON:
        assert tuple(sx[0]) == (1.0, 1.0, 1.0) and tuple(sx[3]) == (9.0, 9.0, 9.0)


# ------------------------------------------------------------------ a6 sort
@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_sort_is_stable_sort(nthreads):
    rng = np.random.default_rng(nthreads)
    keys = rng.integers(0, 1 << 63, size=20000, dtype=np.uint64)
    keys[::7] = keys[3]  # many duplicates: stability matters
    keys[::11] = 0
    keys[5::13] = (1 << 63) - 1
    sk, perm = O.sort(keys, nthreads)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(perm, ref)
    np.testing.assert_array_equal(sk, keys[ref])


# ------------------------------------------------------------------ a7-a9 build
def brute_tree(codes, max_leaf):
    """Top-down: a node splits iff it holds more than max_leaf points and its level is < 21.
    Returns the leaves as (start key, level) in key order and all nodes per level."""
    leaves, levels = [], {}

    def rec(start, level):
        size = 1 << (3 * (21 - level))
        lo, hi = np.searchsorted(codes, np.uint64(start)), np.searchsorted(codes, np.uint64(start + size))
        levels.setdefault(level, []).append((start, int(lo), int(hi)))
        if hi - lo > max_leaf and level < 21:
            for o in range(8):
                rec(start + o * (size >> 3), level + 1)
        else:
            leaves.append((start, level, int(hi - lo)))

    rec(0, 0)
    return leaves, levels


CLOUDS = {
    "clustered": lambda rng: np.vstack([rng.normal(size=(1500, 3)) * 0.3 + [2, 2, 2],
                                        rng.uniform(0, 10, size=(1500, 3)),
                                        np.repeat([[7.0, 1.0, 3.0]], 40, axis=0)]),
    "plane": lambda rng: np.column_stack([rng.uniform(0, 50, 2000), rng.uniform(0, 20, 2000), np.zeros(2000)]),
}


@pytest.mark.parametrize("enc", [O.ENC_MORTON, O.ENC_HILBERT])
@pytest.mark.parametrize("cloud", sorted(CLOUDS))
@pytest.mark.parametrize("max_leaf", [1, 8, 32])
def test_build_matches_closed_form(enc, cloud, max_leaf):
    rng = np.random.default_rng(enc * 1000 + max_leaf + 7 * len(cloud))
    xyz = CLOUDS[cloud](rng)
    sx, keys, perm, c, r = O.sort_points(enc, xyz)
    t = O.Tree(keys, max_leaf, enc, c, r)
    leaves, levels = brute_tree(keys.astype(np.uint64), max_leaf)
    assert t.nLeaf == len(leaves)
    np.testing.assert_array_equal(t.leaves[:-1], np.array([s for s, _, _ in leaves], np.uint64))
    assert int(t.leaves[-1]) == UPPER
    np.testing.assert_array_equal(t.counts, np.array([n for _, _, n in leaves], np.uint64))
    assert t.nInternal == (t.nLeaf - 1) // 7
    assert t.nTotal == t.nLeaf + t.nInternal
    # level-major, key-minor node order; placeholder prefixes; first-child offsets; ranges
    order = [(lv, s, lo, hi) for lv in sorted(levels) for (s, lo, hi) in levels[lv]]
    assert len(order) == t.nTotal
    pref = t.prefixes
    offs = t.offsets
    rngs = t.ranges
    index = {}
    for i, (lv, s, lo, hi) in enumerate(order):
        p = (1 << (3 * lv)) | (s >> (3 * (21 - lv)))
        assert int(pref[i]) == p
        assert (int(rngs[i, 0]), int(rngs[i, 1])) == (lo, hi)
        index[p] = i
    for i, (lv, s, lo, hi) in enumerate(order):
        child = index.get(int(pref[i]) << 3)
        assert int(offs[i]) == (child if child is not None else 0)
    # centres: the cell's midpoint; every point of the node lies in centre +- radii[level]
    cen, rad = t.centers, t.radii
    # halfLengths = 0.5f * eps * (max' - min') with eps = 2^-21 (linear_octree.hpp:443-445,
    # morton_encoder_3d.hpp:26); radii[L] = 2^(21 - L) * halfLengths (:447-453)
    lo3, hl = c - r, 2.0 ** -22 * ((c + r) - (c - r))
    np.testing.assert_array_equal(rad[21], hl)
    for lv in range(22):
        np.testing.assert_array_equal(rad[lv], hl * 2.0 ** (21 - lv))
    for i, (lv, s, lo, hi) in enumerate(order):
        dec = O.hilbert_decode(s) if enc == O.ENC_HILBERT else O.morton_decode(s)
        mask = MAXC ^ ((1 << (21 - lv)) - 1)
        corner = np.array([v & mask for v in dec], np.float64)
        mid = lo3 + (2 * corner + 2.0 ** (21 - lv)) * hl
        np.testing.assert_allclose(cen[i], mid, rtol=1e-12, atol=1e-12)
        if hi > lo:
            p = sx[lo:hi]
            assert (np.abs(p - cen[i]) <= rad[lv] * (1 + 1e-9) + 1e-9).all()


# ------------------------------------------------------------------ a10-a14 radius searches
def brute_inside(xyz, kernel, q, radius):
    """isInside of the four kernels (kernel_3d.hpp:133-138, 238-242; kernel_2d.hpp:59-63, 159-162),
    left-to-right FP64 evaluation without contraction (numpy never fuses)."""
    d1, d2, d3 = xyz[:, 0] - q[0], xyz[:, 1] - q[1], xyz[:, 2] - q[2]
    R = radius * radius
    if kernel == O.K_SPHERE:
        return d1 * d1 + d2 * d2 + d3 * d3 < R
    if kernel == O.K_CIRCLE:
        return d1 * d1 + d2 * d2 <= R
    lo, hi = q - radius, q + radius
    m = (xyz[:, 0] >= lo[0]) & (xyz[:, 0] <= hi[0]) & (xyz[:, 1] >= lo[1]) & (xyz[:, 1] <= hi[1])
    if kernel == O.K_CUBE:
        m &= (xyz[:, 2] >= lo[2]) & (xyz[:, 2] <= hi[2])
    return m


@pytest.fixture(scope="module")
def radius_case():
    rng = np.random.default_rng(123)
    xyz = np.vstack([rng.uniform(0, 20, size=(4000, 3)) * [1, 1, 0.3],
                     rng.normal(size=(1500, 3)) * 0.5 + [5, 5, 2],
                     np.repeat([[12.0, 3.0, 1.0]], 30, axis=0)])
    sx, keys, perm, c, r = O.sort_points(O.ENC_HILBERT, xyz)
    t = O.Tree(keys, 16, O.ENC_HILBERT, c, r)
    qs = np.vstack([sx[rng.integers(0, len(sx), 120)], rng.uniform(-2, 22, size=(40, 3)) * [1, 1, 0.3],
                    [[12.0, 3.0, 1.0]]])
    return sx, t, qs


@pytest.mark.parametrize("kernel", [O.K_SPHERE, O.K_CIRCLE, O.K_CUBE, O.K_SQUARE])
@pytest.mark.parametrize("algo", [O.A_NEIGHBORS, O.A_PRUNE, O.A_STRUCT])
@pytest.mark.parametrize("radius", [0.0, 0.3, 1.7, 40.0])
def test_radius_sets_equal_brute_force(radius_case, kernel, algo, radius):
    sx, t, qs = radius_case
    out, sizes = O.radius(t, sx, kernel, algo, qs, radius)
    cnt = O.radius_count(t, sx, kernel, algo, qs, radius)
    np.testing.assert_array_equal(cnt, sizes)
    for qi, q in enumerate(qs):
        want = np.flatnonzero(brute_inside(sx, kernel, q, radius))
        if algo == O.A_STRUCT:
            rr = out[qi]
            got = np.concatenate([np.arange(a, b) for a, b in rr]) if len(rr) else np.zeros(0, np.int64)
            assert (rr[:, 0] <= rr[:, 1]).all() if len(rr) else True
        else:
            got = out[qi]
            assert len(np.unique(got)) == len(got)
        np.testing.assert_array_equal(np.sort(got), want, err_msg=f"q={qi} k={kernel} a={algo} r={radius}")
        assert sizes[qi] == len(want)


def test_radius_brute_helper_matches_numpy(radius_case):
    sx, t, qs = radius_case
    for kernel in (O.K_SPHERE, O.K_CIRCLE, O.K_CUBE, O.K_SQUARE):
        out, sizes = O.radius_brute(sx, kernel, qs[:30], 1.1)
        for qi in range(30):
            np.testing.assert_array_equal(out[qi], np.flatnonzero(brute_inside(sx, kernel, qs[qi], 1.1)))


def test_radius_kats():
    """test_octree.cpp:63-80, test_octree_advanced.cpp:69-131, test_complex.cpp:37-94."""
    def run(pts, kernel, q, rad, algo=O.A_PRUNE, enc=O.ENC_MORTON, ml=128):
        sx, keys, perm, c, r = O.sort_points(enc, pts)
        t = O.Tree(keys, ml, enc, c, r)
        out, _ = O.radius(t, sx, kernel, algo, [q], rad)
        return sx, out[0]

    oct8 = [[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [10, 10, 10], [10.1, 10, 10], [10, 10.1, 10], [5, 5, 5], [5, 5, 5.1]]
    sx, idx = run(oct8, O.K_SQUARE, [5, 5, 5], 0.25)
    assert len(idx) == 2 and sx[idx[0], 2] == 5.0 and sx[idx[1], 2] == 5.1
    _, idx = run([[0, 0, 0], [0, 0, 0], [1, 1, 1]], O.K_SQUARE, [0, 0, 0], 0.0)
    assert len(idx) == 2
    _, idx = run([[0, 0, 0], [10, 10, 10], [5, 5, 5]], O.K_SQUARE, [10, 10, 10], 0.0)
    assert len(idx) == 1
    _, idx = run([[0, 0, 0], [10, 10, 10], [5, 5, 5]], O.K_SQUARE, [10.1, 10.1, 10.1], 0.01)
    assert len(idx) == 0
    _, idx = run([[1, 1, 1], [2, 0, 0]], O.K_SPHERE, [0, 0, 0], 1.5)
    assert len(idx) == 0
    _, idx = run([[1, 1, 1], [2, 0, 0]], O.K_SPHERE, [0, 0, 0], 2.5)
    assert len(idx) == 2
    sx, idx = run([[1, 1, 1], [2, 2, 2]], O.K_CUBE, [0, 0, 0], 1.5)
    assert len(idx) == 1 and sx[idx[0], 0] == 1.0
    line = [[i * 0.1, 0, 0] for i in range(100)]
    sx, idx = run(line, O.K_SPHERE, [5, 0, 0], 0.55)
    assert len(idx) == len(np.flatnonzero(brute_inside(sx, O.K_SPHERE, np.array([5.0, 0, 0]), 0.55)))
    big = [[1e6, 1e6, 1e6], [1e6 + 0.1, 1e6, 1e6], [1e6, 1e6 + 0.1, 1e6]]
    _, idx = run(big, O.K_SPHERE, [1e6 + 0.05, 1e6 + 0.05, 1e6], 0.1)
    assert len(idx) == 3


def test_neighbor_set_ranges_kat():  # test_neighbor_set.cpp:5-38: ranges [0,3) + [7,10) -> 0,1,2,7,8,9
    pts = [[i, i, i] for i in range(10)]
    sx, keys, perm, c, r = O.sort_points(O.ENC_MORTON, pts)
    t = O.Tree(keys, 2, O.ENC_MORTON, c, r)
    out, sizes = O.radius(t, sx, O.K_CUBE, O.A_STRUCT, [[20, 20, 20]], 1.0)
    assert sizes[0] == 0 and len(out[0]) == 0
    out, sizes = O.radius(t, sx, O.K_CUBE, O.A_STRUCT, [[1, 1, 1]], 1.0)
    got = np.concatenate([np.arange(a, b) for a, b in out[0]])
    assert sizes[0] == 3 and sorted(sx[got, 0].tolist()) == [0.0, 1.0, 2.0]


# ------------------------------------------------------------------ a15 kNN
@pytest.mark.parametrize("k", [1, 5, 50, 6000])
def test_knn_equals_brute_force(radius_case, k):
    sx, t, qs = radius_case
    qs = qs[:40]
    kk = min(k, len(sx))
    idx, d2 = O.knn(t, sx, qs, kk)
    for qi, q in enumerate(qs):
        # the reference's squared distance: (dx*dx + dy*dy) + dz*dz (linear_octree.hpp:844-851)
        dx, dy, dz = sx[:, 0] - q[0], sx[:, 1] - q[1], sx[:, 2] - q[2]
        dd = dx * dx + dy * dy + dz * dz
        order = np.lexsort((np.arange(len(sx)), dd))[:kk]
        np.testing.assert_array_equal(d2[qi], dd[order])
        # the reference's heap (linear_octree.hpp:899-942) orders equal distances arbitrarily: within
        # each distance value the oracle returns a subset of that tie class, all of it below the k-th
        assert len(np.unique(idx[qi])) == kk
        np.testing.assert_array_equal(dd[idx[qi]], d2[qi])
        kth = d2[qi][-1]
        np.testing.assert_array_equal(np.sort(idx[qi][d2[qi] < kth]), np.flatnonzero(dd < kth))


def test_knn_kats():
    """test_octree.cpp:42-61 and test_octree_advanced.cpp:49-67."""
    oct8 = [[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [10, 10, 10], [10.1, 10, 10], [10, 10.1, 10], [5, 5, 5], [5, 5, 5.1]]
    sx, keys, perm, c, r = O.sort_points(O.ENC_MORTON, oct8)
    t = O.Tree(keys, 128, O.ENC_MORTON, c, r)
    idx, d2 = O.knn(t, sx, [[9.9, 9.9, 9.9]], 2)
    assert abs(d2[0, 0] - 0.03) < 1e-4 and sx[idx[0, 0], 0] > 9.0
    pts = [[1, 1, 1], [1, 1, 1], [2, 2, 2]]
    sx, keys, perm, c, r = O.sort_points(O.ENC_MORTON, pts)
    t = O.Tree(keys, 128, O.ENC_MORTON, c, r)
    idx, d2 = O.knn(t, sx, [[1, 1, 1]], 3)
    assert d2[0, 0] == 0 and d2[0, 1] == 0 and d2[0, 2] == 3.0


@pytest.mark.parametrize("kernel", [O.K_CUBE, O.K_SQUARE])
@pytest.mark.parametrize("algo", [O.A_NEIGHBORS, O.A_PRUNE, O.A_STRUCT])
def test_radius_box_vector_radii_equal_brute_force(radius_case, kernel, algo):
    """Vector-radii overloads (kernel_factory.hpp:33-41): box = centre -/+ radii per axis."""
    sx, t, qs = radius_case
    rng = np.random.default_rng(4)
    r3 = rng.uniform(0.0, 3.0, size=(len(qs), 3))
    out, sizes = O.radius_box(t, sx, kernel, algo, qs, r3)
    for qi, q in enumerate(qs):
        lo, hi = q - r3[qi], q + r3[qi]
        m = (sx[:, 0] >= lo[0]) & (sx[:, 0] <= hi[0]) & (sx[:, 1] >= lo[1]) & (sx[:, 1] <= hi[1])
        if kernel == O.K_CUBE:
            m &= (sx[:, 2] >= lo[2]) & (sx[:, 2] <= hi[2])
        want = np.flatnonzero(m)
        rr = out[qi]
        if algo == O.A_STRUCT:
            got = np.concatenate([np.arange(a, b) for a, b in rr]) if len(rr) else np.zeros(0, np.int64)
        else:
            got = rr
        np.testing.assert_array_equal(np.sort(got), want)
    # equal radii reduce to the scalar search
    o1, s1 = O.radius_box(t, sx, kernel, algo, qs, [1.3, 1.3, 1.3])
    o2, s2 = O.radius(t, sx, kernel, algo, qs, 1.3)
    np.testing.assert_array_equal(s1, s2)
    for a, b in zip(o1, o2):
        np.testing.assert_array_equal(a, b)


def test_approx_bounds_and_cutoff():
    """neighborsApprox (linear_octree.hpp:981-1035): reference KAT test_complex.cpp:116-142 (upper >=
    exact >= lower at precision maxDepth / 2), the lower bound's points are a subset of the exact set
    and the exact set of the upper bound's, and with the level at or beyond the tree's depth both
    modes equal neighborsStruct range for range."""
    rng = np.random.default_rng(42)
    xyz = rng.uniform(0.0, 10.0, (1000, 3))
    sx, keys, perm, c, r = O.sort_points(O.ENC_MORTON, xyz)
    t = O.Tree(keys, 32, O.ENC_MORTON, c, r)
    centers = np.vstack([[5.0, 5.0, 5.0], sx[::97]])
    ex, es = O.radius(t, sx, O.K_SPHERE, O.A_STRUCT, centers, 2.0)
    up, us = O.radius_approx(t, sx, O.K_SPHERE, centers, 21 // 2, True, radii=2.0)
    lo, ls = O.radius_approx(t, sx, O.K_SPHERE, centers, 21 // 2, False, radii=2.0)
    assert np.all(us >= es) and np.all(ls <= es)
    expand = lambda rr: set(i for a, b in rr for i in range(int(a), int(b)))  # noqa: E731
    for level in range(0, t.maxDepthSeen + 2):
        up, us = O.radius_approx(t, sx, O.K_SPHERE, centers, level, True, radii=2.0)
        lo, ls = O.radius_approx(t, sx, O.K_SPHERE, centers, level, False, radii=2.0)
        for q in range(len(centers)):
            e = expand(ex[q])
            assert expand(lo[q]) <= e <= expand(up[q]), (level, q)
            if level >= t.maxDepthSeen:
                np.testing.assert_array_equal(up[q], ex[q])
                np.testing.assert_array_equal(lo[q], ex[q])
    # level 0, upper bound: an OVERLAP root adds all 8 children whole
    up, us = O.radius_approx(t, sx, O.K_SPHERE, [[5.0, 5.0, 5.0]], 0, True, radii=2.0)
    assert us[0] == len(sx)


# ------------------------------------------------------------------ LAS ingest (next: data formats)
LAS_CASES = [(minor, fmt) for minor in (2, 3) for fmt in range(6)] + [(4, fmt) for fmt in range(11)]


@pytest.mark.parametrize("minor,fmt", LAS_CASES)
def test_las_oracle_reads_written_fields(tmp_path, minor, fmt):
    """Files packed field by field from the LAS layouts: coordinates X * scale + offset, attributes as
    written (formats 0-5) or through LASlib's legacy fields (6-10), extra bytes and VLRs skipped."""
    from las_files import expected_attributes, write_las
    rng = np.random.default_rng(100 * minor + fmt)
    n = 1500
    X = rng.integers(-2 ** 31, 2 ** 31, n)
    Y = rng.integers(-2 ** 31, 2 ** 31, n)
    Z = rng.integers(-10 ** 6, 10 ** 6, n)
    scale, off = (0.001, 0.01, 0.0001), (652000.25, -4100000.5, 12.0)
    path = str(tmp_path / f"c{minor}_{fmt}.las")
    w = write_las(path, X, Y, Z, fmt, minor=minor, scale=scale, offset=off, extra=3 * (fmt % 2), vlrs=fmt % 3, seed=fmt)
    info, s, o = O.las_header(path)
    assert info[0] == minor and info[1] == fmt and info[3] == n
    assert np.array_equal(s, scale) and np.array_equal(o, off)
    got = O.las_read(path)
    for a, v in enumerate((X, Y, Z)):
        ref = v.astype(np.float64) * scale[a] + off[a]
        assert np.array_equal(got["xyz"][:, a].view(np.uint64), ref.view(np.uint64)), f"axis {a}"
    f, inten, psid, rgb = expected_attributes(w)
    np.testing.assert_array_equal(got["fields"], f)
    np.testing.assert_array_equal(got["intensity"], inten)
    np.testing.assert_array_equal(got["psid"], psid)
    np.testing.assert_array_equal(got["rgb"], rgb)


def test_las_oracle_closed_form_coordinates(tmp_path):
    """Power-of-two scales and integral offsets: every coordinate is exact, X / 4 + 1000 etc."""
    from las_files import write_las
    X = np.arange(-5000, 5000, 7)
    Y = X[::-1] * 3
    Z = X % 251
    path = str(tmp_path / "exact.las")
    write_las(path, X, Y, Z, 1, minor=2, scale=(0.25, 0.5, 0.125), offset=(1000.0, -20.0, 3.0))
    got = O.las_read(path)["xyz"]
    np.testing.assert_array_equal(got[:, 0], X / 4 + 1000.0)
    np.testing.assert_array_equal(got[:, 1], Y / 2 - 20.0)
    np.testing.assert_array_equal(got[:, 2], Z / 8 + 3.0)


def test_las_oracle_rejects(tmp_path):
    from las_files import write_las
    path = str(tmp_path / "z.las")
    write_las(path, [1, 2], [3, 4], [5, 6], 1, compressed=True)
    with pytest.raises(RuntimeError):
        O.las_header(path)
    bad = tmp_path / "bad.las"
    bad.write_bytes(b"LASX" + bytes(400))
    with pytest.raises(RuntimeError):
        O.las_header(str(bad))


# ------------------------------------------------------------------ competitor baseline
@pytest.mark.skipif(not O.nanoflann_available(), reason="oracle/_ref/libref_nanoflann.so not built")
def test_nanoflann_baseline_agrees_with_oracle():
    """The reference's nanoflann KD-tree (its competitor) gives the sphere-kernel counts of the oracle
    (strict < on the same left-to-right sum of squares) and the same k smallest distances."""
    rng = np.random.default_rng(31)
    pts = np.vstack([rng.uniform(0, 30, (6000, 3)) * [1, 1, 0.2], rng.normal(size=(500, 3)) * 0.3 + 5])
    sx, keys, perm, c, r = O.sort_points(O.ENC_HILBERT, pts)
    tree = O.Tree(keys, 32, O.ENC_HILBERT, c, r)
    nf = O.NanoflannTree(sx, 32)
    centers = sx[rng.integers(0, len(sx), 300)]
    for radius in (0.5, 1.5):
        ref = O.radius_count(tree, sx, O.K_SPHERE, O.A_STRUCT, centers, radius)
        np.testing.assert_array_equal(nf.radius_counts(centers, radius), ref)
    bi, bd = O.knn_brute(sx, centers, 12)
    ni, nd = nf.knn(centers, 12)
    np.testing.assert_array_equal(nd, bd)


def test_pointer_octree_competitor_agrees_with_oracle():
    """The reference's pointer octree restated as a competitor (orc_ptree_*): its neighbors() lists
    hold exactly the brute-force sets for all four kernels, its counts equal the linear octree's, its
    leaves hold at most maxPointsLeaf points (the insertion rule of octree.cpp:217-243) and every point
    sits in exactly one leaf; more identical points than a leaf holds stop splitting at depth 64."""
    rng = np.random.default_rng(33)
    pts = np.vstack([rng.uniform(0, 30, (8000, 3)) * [1, 1, 0.2], rng.normal(size=(800, 3)) * 0.3 + 5,
                     np.repeat([[7.0, 7.0, 1.0]], 50, axis=0)])
    sx, keys, perm, c, r = O.sort_points(O.ENC_HILBERT, pts)
    tree = O.Tree(keys, 32, O.ENC_HILBERT, c, r)
    pt = O.PointerOctree(sx, 32, c, r)
    assert pt.n_leaf > 0 and pt.n_internal > 0 and pt.n_leaf == 7 * pt.n_internal + 1
    centers = np.vstack([sx[rng.integers(0, len(sx), 200)], [[7.0, 7.0, 1.0]]])
    for kernel in (O.K_SPHERE, O.K_CUBE, O.K_CIRCLE, O.K_SQUARE):
        for radius in (0.4, 1.5):
            lists = pt.radius(kernel, centers, radius)
            brute, _ = O.radius_brute(sx, kernel, centers, radius)
            for q in range(len(centers)):
                np.testing.assert_array_equal(np.sort(lists[q]), np.sort(brute[q]), err_msg=f"kernel {kernel} q {q}")
            np.testing.assert_array_equal(pt.radius_counts(kernel, centers, radius),
                                          O.radius_count(tree, sx, kernel, O.A_NEIGHBORS, centers, radius))
    # 40 identical points over a leaf size of 8: splitting stops at depth 64, the leaf keeps them all
    dup = np.vstack([rng.uniform(0, 1, (100, 3)), np.repeat([[0.5, 0.5, 0.5]], 40, axis=0)])
    c2, r2 = O.mbb(dup)
    pd = O.PointerOctree(dup, 8, c2, r2)
    assert pd.depth == 64
    got = pd.radius_counts(O.K_SPHERE, np.array([[0.5, 0.5, 0.5]]), 1e-9)
    assert int(got[0]) == 40
