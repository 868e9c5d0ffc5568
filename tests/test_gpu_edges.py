"""GPU edge cases the reference's own tests exercise (empty cloud, duplicates, zero radius, boundary
queries, degenerate axes, large offsets) plus the known answers of those tests, each also checked
bit-exact against the oracle."""
import numpy as np
import pytest

import octree_amd as A
import pyoracle as O

pytestmark = pytest.mark.gpu


def run_both(ctx, pts, enc, ml, kernel, algo, centers, radius):
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    sx, keys, perm, c, r = O.sort_points(enc, pts)
    tree = O.Tree(keys, ml, enc, c, r)
    ref, ref_sizes = O.radius(tree, sx, kernel, algo, centers, radius)
    ctx.set_points(pts)
    gc, gr, gperm, gkeys = ctx.sort_points(enc)
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)
    ctx.build(ml)
    out = []
    for prec in (A.PREC_FP64, A.PREC_FP32):
        counts, lists, _ = ctx.radius(kernel, algo, prec, centers=np.asarray(centers, np.float64), radius=radius,
                                      lists=True)
        np.testing.assert_array_equal(counts, ref_sizes)
        for q in range(len(lists)):
            np.testing.assert_array_equal(lists[q], ref[q])
        out.append(lists)
    check_counts(ctx, kernel, algo, centers, radius, ref_sizes)
    return sx, out[0]


def check_counts(ctx, kernel, algo, centers, radius, ref_sizes):
    """Every count-mode variant (tile shapes, leaf remapping on and off, both precisions) on the same case."""
    try:
        for remap in (None, False):
            ctx.set_remap(remap)
            for g in (0, 16, 32, 64):
                ctx.set_group(g)
                for prec in (A.PREC_FP64, A.PREC_FP32):
                    counts, _ = ctx.radius(kernel, algo, prec, centers=np.asarray(centers, np.float64), radius=radius)
                    np.testing.assert_array_equal(counts, ref_sizes, err_msg=f"remap={remap} G={g} prec={prec}")
    finally:
        ctx.set_group(-1)
        ctx.set_remap(None)


OCT8 = [[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [10, 10, 10], [10.1, 10, 10], [10, 10.1, 10], [5, 5, 5], [5, 5, 5.1]]


def test_kat_square_radius(ctx):  # test_octree.cpp:63-80
    sx, lists = run_both(ctx, OCT8, A.ENC_MORTON, 128, A.K_SQUARE, A.A_PRUNE, [[5, 5, 5]], 0.25)
    assert len(lists[0]) == 2
    assert sx[lists[0][0], 2] == 5.0 and sx[lists[0][1], 2] == 5.1


def test_kat_sphere_and_cube(ctx):  # test_octree_advanced.cpp:103-131
    _, l1 = run_both(ctx, [[1, 1, 1], [2, 0, 0]], A.ENC_MORTON, 128, A.K_SPHERE, A.A_PRUNE, [[0, 0, 0]], 1.5)
    assert len(l1[0]) == 0
    _, l2 = run_both(ctx, [[1, 1, 1], [2, 0, 0]], A.ENC_MORTON, 128, A.K_SPHERE, A.A_PRUNE, [[0, 0, 0]], 2.5)
    assert len(l2[0]) == 2
    sx, l3 = run_both(ctx, [[1, 1, 1], [2, 2, 2]], A.ENC_MORTON, 128, A.K_CUBE, A.A_PRUNE, [[0, 0, 0]], 1.5)
    assert len(l3[0]) == 1 and sx[l3[0][0], 0] == 1.0


def test_kat_zero_radius_and_boundary(ctx):  # test_octree_advanced.cpp:69-101
    sx, l = run_both(ctx, [[0, 0, 0], [0, 0, 0], [1, 1, 1]], A.ENC_MORTON, 128, A.K_SQUARE, A.A_PRUNE, [[0, 0, 0]], 0.0)
    assert len(l[0]) == 2
    _, l = run_both(ctx, [[0, 0, 0], [10, 10, 10], [5, 5, 5]], A.ENC_MORTON, 128, A.K_SQUARE, A.A_PRUNE,
                    [[10, 10, 10], [10.1, 10.1, 10.1]], 0.0)
    assert len(l[0]) == 1 and len(l[1]) == 0


def test_degenerate_line_and_plane(ctx):  # test_complex.cpp:37-75 (zero-extent axes)
    line = [[i * 0.1, 0.0, 0.0] for i in range(100)]
    for enc in (A.ENC_MORTON, A.ENC_HILBERT):
        _, l = run_both(ctx, line, enc, 128, A.K_SPHERE, A.A_PRUNE, [[5.0, 0.0, 0.0]], 0.55)
        assert len(l[0]) == 11
        plane = [[i * 0.1, j * 0.1, 0.0] for i in range(10) for j in range(10)]
        run_both(ctx, plane, enc, 4, A.K_SPHERE, A.A_STRUCT, [[0.5, 0.5, 0.0]], 0.2)


def test_large_coordinates(ctx):  # test_complex.cpp:77-94
    o = 1e6
    _, l = run_both(ctx, [[o, o, o], [o + 0.1, o, o], [o, o + 0.1, o]], A.ENC_MORTON, 128, A.K_SPHERE, A.A_PRUNE,
                    [[o + 0.05, o + 0.05, o]], 0.1)
    assert len(l[0]) == 3


def test_huge_extent_and_radius(ctx):
    """Coordinates and radii near 1e20: FP32 terms of the leaf tests would overflow, so those
    leaf/query pairs must be decided in FP64."""
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1e20, 1e20, size=(3000, 3))
    centers = np.vstack([pts[:40], rng.uniform(-1e20, 1e20, size=(40, 3))])
    for kernel in (A.K_SPHERE, A.K_CIRCLE):
        run_both(ctx, pts, A.ENC_HILBERT, 16, kernel, A.A_STRUCT, centers, 4e19)


def test_duplicates_force_max_depth(ctx):
    pts = np.vstack([np.tile([[1.0, 2.0, 3.0]], (300, 1)), [[0, 0, 0], [9, 9, 9]]])
    for enc in (A.ENC_MORTON, A.ENC_HILBERT):
        for algo in (A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT):
            run_both(ctx, pts, enc, 32, A.K_SPHERE, algo, [[1, 2, 3], [0, 0, 0]], 0.5)
    assert ctx.info["max_depth_seen"] == 21


def test_empty_cloud(ctx):  # test_octree_advanced.cpp:33-47
    ctx.set_points(np.zeros((0, 3)))
    c, r, perm, keys = ctx.sort_points(A.ENC_MORTON)
    assert np.all(np.isneginf(c)) and np.all(np.isneginf(r))
    info = ctx.build(128)
    assert info["n_leaf"] == 1 and info["n_total"] == 1
    counts, lists, _ = ctx.radius(A.K_SQUARE, A.A_PRUNE, A.PREC_FP64, centers=[[0.0, 0.0, 0.0]], radius=1.0, lists=True)
    assert counts[0] == 0 and len(lists[0]) == 0
    gi, gd, _ = ctx.knn(1, centers=[[0.0, 0.0, 0.0]])
    assert gi[0, 0] == -1 and np.isnan(gd[0, 0])


def test_knn_duplicates_and_k_above_n(ctx):  # test_octree_advanced.cpp:49-67
    pts = np.array([[1, 1, 1], [1, 1, 1], [2, 2, 2]], np.float64)
    ctx.set_points(pts)
    ctx.sort_points(A.ENC_MORTON)
    ctx.build(128)
    gi, gd, _ = ctx.knn(5, centers=[[1.0, 1.0, 1.0]])
    assert list(gd[0, :3]) == [0.0, 0.0, 3.0] and list(gi[0, 3:]) == [-1, -1]


@pytest.mark.parametrize("offset", [0.0, 1e6, 1e12])
def test_knn_offsets_and_far_queries(ctx, offset):
    """kNN with large coordinate offsets (the pruning margin scales with the magnitudes) and with
    centres outside the cloud, both precisions, against brute force."""
    print(f"[gpu] knn offset={offset}", flush=True)
    ctx.current = None
    xyz = A.uniform_box(20_000, seed=3) * 0.01 + offset
    ctx.set_points(xyz)
    _, _, perm, _ = ctx.sort_points(A.ENC_HILBERT)
    ctx.build(16)
    sx = xyz[perm.astype(np.int64)]
    rng = np.random.default_rng(4)
    centers = np.vstack([sx[rng.integers(0, len(sx), 200)],
                         offset + rng.uniform(-2.0, 3.0, (50, 3)),
                         [[offset - 1e3, offset, offset], [offset, offset, offset + 7.0]]])
    bi, bd = O.knn_brute(sx, centers, 12)
    for prec in (A.PREC_FP64, A.PREC_FP32):
        gi, gd, _ = ctx.knn(12, precision=prec, centers=centers)
        np.testing.assert_array_equal(gi.astype(np.int64), bi)
        assert np.array_equal(gd.view(np.uint64), bd.view(np.uint64))


def test_single_point_and_small_leaf(ctx):
    for n in (1, 2, 7, 33, 129, 1000):
        pts = A.uniform_box(n, seed=n)
        run_both(ctx, pts, A.ENC_HILBERT, 4, A.K_CIRCLE, A.A_STRUCT, pts[: min(n, 50)], 7.5)


def test_bad_arguments(ctx):
    ctx.set_points(A.uniform_box(1000, seed=3))
    with pytest.raises(A.OctError):
        ctx.encode(3, np.zeros(3), np.ones(3))  # 2-D encoders cannot build a LinearOctree
    ctx.sort_points(A.ENC_MORTON)
    ctx.build(16)
    with pytest.raises(A.OctError):
        ctx.radius(7, A.A_PRUNE, A.PREC_FP64, centers=[[0.0, 0.0, 0.0]], radius=1.0)
    with pytest.raises(A.OctError):
        ctx.radius(A.K_SPHERE, A.A_PRUNE, A.PREC_FP64, idx=np.array([5000], np.uint32), radius=1.0)


def test_point_struct_layout(ctx):
    """32-byte Point rows (x, y, z, id) give the same keys as packed xyz (point.hpp:23-29)."""
    xyz = A.uniform_box(5000, seed=8)
    rows = np.concatenate([xyz, np.arange(5000, dtype=np.float64)[:, None]], axis=1)
    ctx.set_points(xyz)
    _, _, p1, k1 = ctx.sort_points(A.ENC_HILBERT)
    ctx.set_points(rows)
    _, _, p2, k2 = ctx.sort_points(A.ENC_HILBERT)
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_array_equal(p1, p2)


def test_key_codec_matches_oracle(ctx):
    """PointEncoder::encode / decode of single coordinates, batched on the device."""
    rng = np.random.default_rng(21)
    m = (1 << 21) - 1
    xyz = np.vstack([rng.integers(0, 1 << 21, size=(5000, 3)), [[0, 0, 0], [m, m, m], [1000, 2000, 3000]]]).astype(np.uint32)
    for enc, fe, fd in ((A.ENC_MORTON, O.morton_encode, O.morton_decode), (A.ENC_HILBERT, O.hilbert_encode, O.hilbert_decode)):
        keys = ctx.encode_coords(enc, xyz)
        want = np.array([fe(*t) for t in xyz], np.uint64)
        np.testing.assert_array_equal(keys, want)
        np.testing.assert_array_equal(ctx.decode_keys(enc, keys), xyz)


def test_radius_box_vector_radii(ctx):
    """Cube / square searches with per-axis half extents (Vector overloads) against the oracle."""
    rng = np.random.default_rng(8)
    pts = np.vstack([rng.uniform(0, 30, size=(20000, 3)) * [1, 1, 0.2], rng.normal(size=(3000, 3)) + [10, 10, 3]])
    sx, keys, perm, c, r = O.sort_points(A.ENC_HILBERT, pts)
    tree = O.Tree(keys, 32, A.ENC_HILBERT, c, r)
    ctx.set_points(pts)
    ctx.sort_points(A.ENC_HILBERT)
    ctx.build(32)
    centers = np.vstack([sx[rng.integers(0, len(sx), 300)], rng.uniform(-1, 31, size=(50, 3))])
    r3q = rng.uniform(0, 2.5, size=(len(centers), 3))
    for kernel in (A.K_CUBE, A.K_SQUARE):
        for algo in (A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT):
            for r3 in (r3q, np.array([0.7, 1.9, 0.4])):
                ref, ref_sizes = O.radius_box(tree, sx, kernel, algo, centers, r3)
                cnt, _ = ctx.radius_box(kernel, algo, r3, centers=centers)
                np.testing.assert_array_equal(cnt, ref_sizes)
                cnt, lists, _ = ctx.radius_box(kernel, algo, r3, centers=centers, lists=True)
                for q in range(len(centers)):
                    np.testing.assert_array_equal(lists[q], ref[q])
    with pytest.raises(A.OctError):
        ctx.radius_box(A.K_SPHERE, A.A_PRUNE, [1.0, 1.0, 1.0], centers=centers)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 5, 21])
def test_radius_approx_levels(ctx, level):
    """neighborsApprox counts and Struct range lists against the oracle, both bound modes."""
    print(f"[gpu] approx level={level}", flush=True)
    ctx.current = None
    xyz = A.lidar_tile(60_000, seed=8, extent=50.0)
    sx, keys, perm, c, r = O.sort_points(O.ENC_HILBERT, xyz)
    tree = O.Tree(keys, 32, O.ENC_HILBERT, c, r)
    ctx.set_points(xyz)
    ctx.sort_points(A.ENC_HILBERT)
    ctx.build(32)
    idx = A.random_indices(400, len(xyz), seed=level)
    centers = sx[idx.astype(np.int64)]
    for upper in (True, False):
        for kernel in (A.K_SPHERE, A.K_CIRCLE, A.K_CUBE, A.K_SQUARE):
            ref, ref_sizes = O.radius_approx(tree, sx, kernel, centers, level, upper, radii=1.7)
            for prec in (A.PREC_FP64, A.PREC_FP32):
                counts, lists, _ = ctx.radius_approx(kernel, prec, level, upper, idx=idx, radius=1.7, lists=True)
                np.testing.assert_array_equal(counts, ref_sizes, err_msg=f"k={kernel} upper={upper} prec={prec}")
                for q in range(len(idx)):
                    np.testing.assert_array_equal(lists[q], ref[q])
                cnt_only, _ = ctx.radius_approx(kernel, prec, level, upper, idx=idx, radius=1.7)
                np.testing.assert_array_equal(cnt_only, ref_sizes)
        ref, ref_sizes = O.radius_approx(tree, sx, A.K_CUBE, centers, level, upper, radii3=[1.0, 2.0, 0.5])
        counts, lists, _ = ctx.radius_approx(A.K_CUBE, A.PREC_FP64, level, upper, idx=idx, radius3=[1.0, 2.0, 0.5], lists=True)
        np.testing.assert_array_equal(counts, ref_sizes)
