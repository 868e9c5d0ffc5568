"""The benchmark's workload configurations at their stated sizes, on the GPU, against the oracle.

C1  1 M uniform points (100 x 100 x 20 m, 5 pts/m^3), Morton, maxLeaf 32, 10 000 random centres
    (with replacement), r = 2.5 m: the result lists of the three search algorithms x four kernels in
    both precisions equal the oracle's, entry for entry.
C2  the 50 M-point airborne tile (1 km^2, Hilbert, maxLeaf 32 and 128): keys, permutation and every
    tree array bit-exact against the oracle; all-points counts for r in {1, 2.5, 5} m x {sphere,
    cube, square, circle}, FP32 + recheck equal to FP64 for every one of the 50 M queries, equal to
    the oracle on a strided sample, Struct range lists equal to the oracle's on a smaller one; the
    kernels' traversal counters (points tested, nodes classified) equal the oracle's per-query
    statistics of the reference traversal.
C3  the same tile, every point a kNN query, k in {10, 50, 250, 1000}, in chunks of queries: FP32 +
    recheck equal to FP64 for every query (indices and distances, compared by a device checksum), squared
    distances equal to the oracle's best-first search on a strided sample, index lists equal to a
    brute force over a crop around the query on a few queries.
The 256-cloud batch (C4) is in test_gpu_batch_c4.py, the 400 M-point sweep (C5) in test_gpu_scale.py.
"""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from crop_knn import crop_brute_knn

pytestmark = pytest.mark.gpu

N50 = 50_000_000
KERNELS = (A.K_SPHERE, A.K_CUBE, A.K_SQUARE, A.K_CIRCLE)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_tree(ctx, tree):
    g = ctx.get_tree()
    assert ctx.info["n_leaf"] == tree.nLeaf and ctx.info["n_total"] == tree.nTotal
    assert ctx.info["max_depth_seen"] == tree.maxDepthSeen
    np.testing.assert_array_equal(g["prefixes"], tree.prefixes)
    np.testing.assert_array_equal(g["offsets"], tree.offsets[:tree.nTotal])
    np.testing.assert_array_equal(g["ranges"], tree.ranges)
    assert np.array_equal(g["centers"].view(np.uint64), tree.centers.view(np.uint64)), "centres not bit-exact"


# ------------------------------------------------------------------ C1
@pytest.fixture(scope="module")
def c1():
    xyz = A.uniform_box(1_000_000, seed=1)
    sx, keys, perm, c, r = O.sort_points(A.ENC_MORTON, xyz)
    tree = O.Tree(keys, 32, A.ENC_MORTON, c, r)
    ctx = A.Octree(0)
    ctx.set_points(xyz)
    _, _, gperm, gkeys = ctx.sort_points(A.ENC_MORTON)
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)
    ctx.build(32)
    check_tree(ctx, tree)
    idx = A.random_indices(10_000, len(xyz), seed=1)
    yield ctx, sx, tree, idx
    ctx.close()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("algo", [A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT])
def test_c1_lists(c1, kernel, algo):
    ctx, sx, tree, idx = c1
    ref, sizes = O.radius(tree, sx, kernel, algo, sx[idx.astype(np.int64)], 2.5)
    for prec in (A.PREC_FP64, A.PREC_FP32):
        say(f"[gpu] C1 lists kernel={kernel} algo={algo} prec={prec}")
        counts, lists, _ = ctx.radius(kernel, algo, prec, idx=idx, radius=2.5, lists=True)
        np.testing.assert_array_equal(counts, sizes)
        for q in range(len(idx)):
            if not np.array_equal(lists[q], ref[q]):
                raise AssertionError(f"C1 query {q} (point {idx[q]}) kernel {kernel} algo {algo} prec {prec}: lists differ")


# ------------------------------------------------------------------ C2 / C3
@pytest.fixture(scope="module")
def tile50():
    say("[gpu] C2 tile: 50 M points")
    xyz = A.lidar_tile(N50, seed=2, extent=1000.0)
    sx, keys, perm, c, r = O.sort_points(A.ENC_HILBERT, xyz)
    trees = {ml: O.Tree(keys, ml, A.ENC_HILBERT, c, r) for ml in (32, 128)}
    ctx = A.Octree(0)
    ctx.set_points(xyz)
    del xyz
    _, _, gperm, gkeys = ctx.sort_points(A.ENC_HILBERT)
    state = {"ctx": ctx, "sx": sx, "trees": trees, "keys_ok": bool(np.array_equal(gkeys, keys)),
             "perm_ok": bool(np.array_equal(gperm, perm)), "ml": None}
    del gperm, gkeys, keys, perm
    yield state
    ctx.close()


def use_tree(t, ml):
    if t["ml"] != ml:
        say(f"[gpu] C2 build maxLeaf {ml}")
        t["ctx"].build(ml)
        t["ml"] = ml
    return t["ctx"], t["sx"], t["trees"][ml]


def test_c2_keys_and_permutation(tile50):
    assert tile50["keys_ok"], "sorted keys differ from the oracle"
    assert tile50["perm_ok"], "permutation differs from the oracle"


@pytest.mark.parametrize("ml", [32, 128])
def test_c2_tree(tile50, ml):
    ctx, sx, tree = use_tree(tile50, ml)
    check_tree(ctx, tree)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("radius", [1.0, 2.5, 5.0])
def test_c2_all_points_counts(tile50, radius, kernel):
    """Every point a query (the harness' -s all mode), neighborsStruct counts."""
    ctx, sx, tree = use_tree(tile50, 128)
    say(f"[gpu] C2 all-points counts r={radius} kernel={kernel}")
    c32, s32 = ctx.radius(kernel, A.A_STRUCT, A.PREC_FP32, radius=radius, nq=N50)
    c64, s64 = ctx.radius(kernel, A.A_STRUCT, A.PREC_FP64, radius=radius, nq=N50)
    say(f"      FP32 {s32['kernel_ms']:.1f} ms, FP64 {s64['kernel_ms']:.1f} ms, mean {c64.mean():.1f}")
    assert np.array_equal(c32, c64), f"FP32 + recheck differs from FP64 on {int(np.count_nonzero(c32 != c64))} queries"
    assert s32["total_neighbors"] == s64["total_neighbors"] == int(c64.astype(np.int64).sum())
    samp = np.arange(7, N50, N50 // 4000, dtype=np.int64)
    ref = O.radius_count(tree, sx, kernel, A.A_STRUCT, sx[samp], radius)
    np.testing.assert_array_equal(c64[samp], ref)


@pytest.mark.parametrize("algo", [A.A_NEIGHBORS, A.A_PRUNE])
def test_c2_other_algorithms(tile50, algo):
    ctx, sx, tree = use_tree(tile50, 128)
    for kernel in (A.K_SPHERE, A.K_CUBE):
        say(f"[gpu] C2 all-points counts algo={algo} kernel={kernel}")
        c32, _ = ctx.radius(kernel, algo, A.PREC_FP32, radius=2.5, nq=N50)
        c64, _ = ctx.radius(kernel, algo, A.PREC_FP64, radius=2.5, nq=N50)
        assert np.array_equal(c32, c64)
        samp = np.arange(3, N50, N50 // 2000, dtype=np.int64)
        np.testing.assert_array_equal(c64[samp], O.radius_count(tree, sx, kernel, algo, sx[samp], 2.5))


@pytest.mark.parametrize("radius", [1.0, 2.5, 5.0])
def test_c2_struct_lists_sample(tile50, radius):
    ctx, sx, tree = use_tree(tile50, 128)
    samp = np.arange(11, N50, N50 // 500, dtype=np.int64)
    for kernel in KERNELS:
        ref, sizes = O.radius(tree, sx, kernel, A.A_STRUCT, sx[samp], radius)
        for prec in (A.PREC_FP64, A.PREC_FP32):
            say(f"[gpu] C2 lists r={radius} kernel={kernel} prec={prec}")
            counts, lists, _ = ctx.radius(kernel, A.A_STRUCT, prec, idx=samp.astype(np.uint32), radius=radius, lists=True)
            np.testing.assert_array_equal(counts, sizes)
            for q in range(len(samp)):
                assert np.array_equal(lists[q], ref[q]), f"query {samp[q]}: lists differ"


@pytest.mark.parametrize("group", [0, 16, 32, 64])
def test_c2_traversal_counters_equal_oracle(tile50, group):
    """points_tested and nodes_classified (the roofline's work measure) equal the reference
    traversal's per-query statistics, summed over a strided sample, for every tile shape."""
    ctx, sx, tree = use_tree(tile50, 128)
    samp = np.arange(5, N50, N50 // 3000, dtype=np.int64)
    try:
        ctx.set_group(group)
        for kernel, algo in ((A.K_SPHERE, A.A_STRUCT), (A.K_SPHERE, A.A_PRUNE), (A.K_CUBE, A.A_NEIGHBORS),
                             (A.K_CIRCLE, A.A_STRUCT)):
            _, sizes, st = O.radius(tree, sx, kernel, algo, sx[samp], 2.5, stats=True)
            tested, classified = int(st[:, 0].sum()), int(st[:, 1].sum())
            for prec in (A.PREC_FP64, A.PREC_FP32):
                counts, g = ctx.radius(kernel, algo, prec, idx=samp.astype(np.uint32), radius=2.5)
                np.testing.assert_array_equal(counts, sizes)
                assert g["points_tested"] == tested, (group, kernel, algo, prec, g["points_tested"], tested)
                assert g["nodes_classified"] == classified, (group, kernel, algo, prec, g["nodes_classified"], classified)
    finally:
        ctx.set_group(-1)


@pytest.mark.parametrize("k", [10, 50, 250, 1000])
def test_c3_all_points_knn(tile50, k):
    """Every point a query, in chunks kept on the device: FP32 + recheck and FP64 results compared by a
    device checksum of every (index, FP32 d^2) entry, every row ascending with indices in the cloud;
    rows of a strided sample fetched and compared with the oracle."""
    ctx, sx, tree = use_tree(tile50, 128)
    chunk = min(N50, max(1, (1 << 30) // k))
    samp = np.arange(13, N50, N50 // 2000, dtype=np.int64)
    got = {A.PREC_FP32: ([], []), A.PREC_FP64: ([], [])}
    mismatched = 0
    for first in range(0, N50, chunk):
        nq = min(chunk, N50 - first)
        sel = samp[(samp >= first) & (samp < first + nq)] - first
        sums = {}
        for prec in (A.PREC_FP32, A.PREC_FP64):
            ctx.knn_dev(k, prec, first=first, nq=nq)
            sums[prec] = ctx.knn_checksum()
            assert sums[prec][1] == 0, f"chunk {first} prec {prec}: {sums[prec][1]} rows not ascending or out of range"
            for q in sel:
                oi, od = ctx.knn_fetch(int(q), 1)
                got[prec][0].append(oi[0])
                got[prec][1].append(od[0])
        same = sums[A.PREC_FP32] == sums[A.PREC_FP64]
        mismatched += not same
        say(f"[gpu] C3 k={k} queries {first}..{first + nq} equal={same}")
    assert mismatched == 0, f"FP32 + recheck differs from FP64 in {mismatched} chunks"
    oi, od = O.knn(tree, sx, sx[samp], k)  # the reference's best-first search
    for prec in (A.PREC_FP32, A.PREC_FP64):
        gd = np.vstack(got[prec][1])
        assert np.array_equal(gd.view(np.uint32), od.astype(np.float32).view(np.uint32)), f"prec {prec}: distances"
    gi = np.vstack(got[A.PREC_FP64][0]).astype(np.int64)
    for row in (0, 1, len(samp) // 2, len(samp) - 1):
        bi, bd = crop_brute_knn(sx, tree, sx[samp[row]], k, float(od[row, -1]))
        np.testing.assert_array_equal(gi[row], bi, err_msg=f"query {samp[row]}")
        assert np.array_equal(bd, od[row])
