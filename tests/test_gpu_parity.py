"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on identical inputs.

Bar: bit-exact for keys, permutation, tree arrays (including FP64 centres), neighbour lists (same
entries in the same order as the reference's emission order), kNN indices and FP64 distances.
"""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_pipeline(xyz, enc, max_leaf):
    sx, keys, perm, c, r = O.sort_points(enc, xyz)
    tree = O.Tree(keys, max_leaf, enc, c, r)
    return sx, keys, perm, c, r, tree


def gpu_pipeline(ctx, xyz, enc, max_leaf, tag=None):
    ctx.set_points(xyz)
    c, r, perm, keys = ctx.sort_points(enc)
    info = ctx.build(max_leaf)
    ctx.current = tag
    return c, r, perm, keys, info


def ensure(ctx, case):
    """Rebuild the case's GPU tree only if another test replaced it."""
    name, xyz, enc, ml = case[:4]
    if getattr(ctx, "current", None) != name:
        say(f"[gpu] rebuild {name}")
        gpu_pipeline(ctx, xyz, enc, ml, tag=name)


def check_tree(ctx, tree):
    g = ctx.get_tree()
    assert ctx.info["n_leaf"] == tree.nLeaf
    assert ctx.info["n_internal"] == tree.nInternal
    assert ctx.info["n_total"] == tree.nTotal
    assert ctx.info["max_depth_seen"] == tree.maxDepthSeen
    np.testing.assert_array_equal(g["prefixes"], tree.prefixes)
    np.testing.assert_array_equal(g["offsets"], tree.offsets[:tree.nTotal])
    np.testing.assert_array_equal(g["ranges"], tree.ranges)
    assert np.array_equal(g["centers"].view(np.uint64), tree.centers.view(np.uint64)), "centres not bit-exact"
    assert np.array_equal(g["radii"].view(np.uint64), tree.radii.view(np.uint64))


CASES = {
    "uniform_1M_morton_32": lambda: (A.uniform_box(1_000_000, seed=1), A.ENC_MORTON, 32),
    "tile_200m_hilbert_32": lambda: (A.lidar_tile(2_000_000, seed=2, extent=200.0), A.ENC_HILBERT, 32),
    "tile_200m_hilbert_128": lambda: (A.lidar_tile(2_000_000, seed=2, extent=200.0), A.ENC_HILBERT, 128),
    "tile_100m_morton_128": lambda: (A.lidar_tile(500_000, seed=3, extent=100.0), A.ENC_MORTON, 128),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, ctx):
    name = request.param
    say(f"[case] {name}")
    xyz, enc, ml = CASES[name]()
    ref = oracle_pipeline(xyz, enc, ml)
    say(f"[gpu] pipeline {name} n={len(xyz)}")
    got = gpu_pipeline(ctx, xyz, enc, ml, tag=name)
    return name, xyz, enc, ml, ref, got


def test_bbox_keys_perm(case, ctx):
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), (gc, gr, gperm, gkeys, info) = case
    assert np.array_equal(gc.view(np.uint64), c.view(np.uint64))
    assert np.array_equal(gr.view(np.uint64), r.view(np.uint64))
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)


def test_unsorted_keys(case, ctx):
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    say(f"[gpu] encode {name}")
    ctx.set_points(xyz)
    ctx.current = None
    k = ctx.encode(enc, c, r)
    np.testing.assert_array_equal(k, O.encode(enc, xyz, c, r))


def test_tree(case, ctx):
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    say(f"[gpu] tree {name}")
    ensure(ctx, case)
    check_tree(ctx, tree)


RADII = {"uniform_1M_morton_32": [2.5], "tile_200m_hilbert_32": [1.0, 2.5, 5.0],
         "tile_200m_hilbert_128": [2.5], "tile_100m_morton_128": [0.75, 2.5]}


@pytest.mark.parametrize("kernel", [A.K_SPHERE, A.K_CIRCLE, A.K_CUBE, A.K_SQUARE])
@pytest.mark.parametrize("algo", [A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT])
def test_radius_lists(case, ctx, kernel, algo):
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    ensure(ctx, case)
    n = len(xyz)
    idx = A.random_indices(4000, n, seed=11)
    for rad in RADII[name]:
        ref, ref_sizes = O.radius(tree, sx, kernel, algo, sx[idx.astype(np.int64)], rad)
        for prec in (A.PREC_FP64, A.PREC_FP32):
            say(f"[gpu] radius {name} kernel={kernel} algo={algo} r={rad} prec={prec}")
            counts, lists, st = ctx.radius(kernel, algo, prec, idx=idx, radius=rad, lists=True)
            say(f"      kernel {st['kernel_ms']:.2f} ms, rechecked {st['rechecked']}")
            np.testing.assert_array_equal(counts, ref_sizes)
            for q in range(len(idx)):
                if not np.array_equal(lists[q], ref[q]):
                    raise AssertionError(f"query {q} (point {idx[q]}): lists differ\n gpu={lists[q][:20]}\n ref={ref[q][:20]}")


@pytest.mark.parametrize("group", [0, 16, 32, 64])
def test_radius_counts_every_group(case, ctx, group):
    """Count-mode kernels (one lane per query, or G queries per wave; leaf remapping on and off)
    against the oracle sizes."""
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    ensure(ctx, case)
    idx = A.random_indices(3000, len(xyz), seed=17)
    ctx.set_group(group)
    try:
        for kernel in (A.K_SPHERE, A.K_CIRCLE, A.K_CUBE, A.K_SQUARE):
            for algo in (A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT):
                ref = O.radius_count(tree, sx, kernel, algo, sx[idx.astype(np.int64)], RADII[name][-1])
                for remap in (None, False):
                    ctx.set_remap(remap)
                    for prec in (A.PREC_FP64, A.PREC_FP32):
                        counts, st = ctx.radius(kernel, algo, prec, idx=idx, radius=RADII[name][-1])
                        np.testing.assert_array_equal(counts, ref, err_msg=f"G={group} k={kernel} a={algo} p={prec} remap={remap}")
    finally:
        ctx.set_group(-1)
        ctx.set_remap(None)


def test_matrix_core_knob_refused(ctx):
    """The point tests run on the vector ALUs: the matrix-core switch accepts only off / default."""
    ctx.set_mfma(False)
    ctx.set_mfma(None)
    with pytest.raises(A.OctError):
        ctx.set_mfma(True)


@pytest.mark.parametrize("prec", [A.PREC_FP64, A.PREC_FP32])
def test_radius_all_points_counts(case, ctx, prec):
    """All-points counts (the harness' -s all mode) against the oracle on a strided sample."""
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    ensure(ctx, case)
    n = len(xyz)
    say(f"[gpu] all-points {name} prec={prec}")
    counts, st = ctx.radius(A.K_SPHERE, A.A_STRUCT, prec, radius=2.5, nq=n)
    sample = np.arange(0, n, 97)
    ref = O.radius_count(tree, sx, O.K_SPHERE, O.A_STRUCT, sx[sample], 2.5)
    np.testing.assert_array_equal(counts[sample], ref)
    assert st["total_neighbors"] == int(counts.astype(np.int64).sum())


def test_radius_dev_index_buffer(case, ctx):
    """oct_radius_dev with a device buffer of sorted-cloud indices as the centres (one SFC range of
    the queries, as a GPU's share of a multi-GPU search): its neighbour total equals the host-index
    entry point's counts over the same centres, which equal the oracle's."""
    name, xyz, enc, ml, (sx, keys, perm, c, r, tree), _ = case
    ensure(ctx, case)
    n = len(xyz)
    q0, nq = n // 3, min(20_000, n - n // 3)
    idx = np.arange(q0, q0 + nq, dtype=np.uint32)
    say(f"[gpu] device index buffer {name} [{q0}, {q0 + nq})")
    counts, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, idx=idx, radius=2.5)
    ref = O.radius_count(tree, sx, O.K_SPHERE, O.A_STRUCT, sx[idx.astype(np.int64)], 2.5)
    np.testing.assert_array_equal(counts, ref)
    buf = A.DeviceArray(np.arange(n, dtype=np.uint32))
    try:
        for prec in (A.PREC_FP32, A.PREC_FP64):
            st = ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, 2.5, nq=nq, stats=True, idx_ptr=buf.ptr(q0))
            assert st["total_neighbors"] == int(ref.astype(np.int64).sum())
    finally:
        buf.close()


_KNN_CASE = {}


def knn_case(ctx):
    """60 k-point tile (Hilbert, maxLeaf 32) with its oracle pipeline, built on the GPU once."""
    if not _KNN_CASE:
        xyz = A.lidar_tile(60_000, seed=5, extent=40.0)
        _KNN_CASE["xyz"] = xyz
        _KNN_CASE["oracle"] = oracle_pipeline(xyz, A.ENC_HILBERT, 32)
    if getattr(ctx, "current", None) != "knn":
        gpu_pipeline(ctx, _KNN_CASE["xyz"], A.ENC_HILBERT, 32, tag="knn")
    return _KNN_CASE["oracle"]


@pytest.mark.parametrize("prec", [A.PREC_FP64, A.PREC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("k", [1, 10, 17, 50, 100, 250, 400, 600, 1000])
def test_knn_exact(ctx, k, prec):
    say(f"[gpu] knn k={k} prec={prec}")
    sx, keys, perm, c, r, tree = knn_case(ctx)
    idx = A.random_indices(500, len(sx), seed=13)
    gi, gd, st = ctx.knn(k, precision=prec, idx=idx)
    bi, bd = O.knn_brute(sx, sx[idx.astype(np.int64)], k)
    np.testing.assert_array_equal(gi.astype(np.int64), bi)
    assert np.array_equal(gd.view(np.uint64), bd.view(np.uint64))
    assert st["total_neighbors"] == 500 * k
    if k in (10, 250):
        # the reference's best-first order returns the same squared distances
        hi, hd = O.knn(tree, sx, sx[idx.astype(np.int64)], k)
        assert np.array_equal(hd.view(np.uint64), bd.view(np.uint64))


@pytest.mark.parametrize("k,lanes", [(10, 2), (10, 8), (30, 2), (30, 8), (50, 4), (50, 16), (100, 8), (100, 32),
                                     (250, 16), (250, 32)])
def test_knn_every_shape(ctx, k, lanes):
    """Every lanes-per-query shape the kNN kernel offers returns the same exact sets."""
    say(f"[gpu] knn shape k={k} lanes={lanes}")
    sx, keys, perm, c, r, tree = knn_case(ctx)
    idx = A.random_indices(300, len(sx), seed=k + lanes)
    bi, bd = O.knn_brute(sx, sx[idx.astype(np.int64)], k)
    try:
        ctx.set_knn_lanes(lanes)
        for prec in (A.PREC_FP64, A.PREC_FP32):
            gi, gd, _ = ctx.knn(k, precision=prec, idx=idx)
            np.testing.assert_array_equal(gi.astype(np.int64), bi)
            assert np.array_equal(gd.view(np.uint64), bd.view(np.uint64))
    finally:
        ctx.set_knn_lanes(None)


def test_knn_autotune_keeps_results(ctx):
    sx, keys, perm, c, r, tree = knn_case(ctx)
    best, ms = ctx.autotune_knn(20, A.PREC_FP32, sample_nq=20_000)
    assert best in (2, 4, 8) and ctx.tuned_knn(20, A.PREC_FP32) == best
    assert np.all(np.isfinite(ms))
    idx = A.random_indices(300, len(sx), seed=3)
    gi, gd, _ = ctx.knn(20, precision=A.PREC_FP32, idx=idx)
    bi, bd = O.knn_brute(sx, sx[idx.astype(np.int64)], 20)
    np.testing.assert_array_equal(gi.astype(np.int64), bi)
    with pytest.raises(A.OctError):
        ctx.set_knn_lanes(64)
        try:
            ctx.knn(20, precision=A.PREC_FP32, idx=idx)  # 64 lanes is not offered for k <= 32
        finally:
            ctx.set_knn_lanes(None)


@pytest.mark.parametrize("prec", [A.PREC_FP64, A.PREC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("k", [10, 50])
def test_knn_dev_all_points(ctx, k, prec):
    """Device-resident all-points kNN over a chunk of sorted points: int32 indices equal the brute
    force, FP32 distances are the FP64 distances rounded to nearest."""
    say(f"[gpu] knn_dev k={k} prec={prec}")
    sx, keys, perm, c, r, tree = knn_case(ctx)
    first, nq = 20_011, 3_000
    st = ctx.knn_dev(k, precision=prec, first=first, nq=nq, stats=True)
    gi, gf = ctx.knn_fetch()
    assert gi.shape == (nq, k) and st["total_neighbors"] == nq * k
    bi, bd = O.knn_brute(sx, sx[first:first + nq], k)
    np.testing.assert_array_equal(gi.astype(np.int64), bi)
    assert np.array_equal(gf.view(np.uint32), bd.astype(np.float32).view(np.uint32))
    # a slice of the kept results
    si, sf = ctx.knn_fetch(100, 7)
    np.testing.assert_array_equal(si, gi[100:107])
