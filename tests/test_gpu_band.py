"""Adversarial tests of the FP32 screening (FP32_RECHECK) against the oracle's FP64 decisions.

Clouds from band_clouds.py put points exactly on the radius (sphere `<` vs circle `<=`), one ulp on
either side of it, and across the width of the FP32 band; points and radii at magnitudes around the
overflow guard of the leaf-recentred FP32 terms; query centres on the equalities of the node
judgements; and kNN queries whose k-th neighbour ties with others. Every count-mode shape (one lane
per query, 16 / 32 / 64 queries per wave, matrix-core or vector point tests, leaf remapping on and
off), the list kernels and kNN in both precisions must return exactly the oracle's results.

The CPU tests at the end check that the constructions really hit the thresholds they aim at.
"""
import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from band_clouds import (EXACT, boundary_cloud, d2_circle, d2_sphere, lattice_cloud, magnitude_cloud,
                         node_boundary_cloud, node_boundary_queries)

KERNELS = (A.K_SPHERE, A.K_CIRCLE, A.K_CUBE, A.K_SQUARE)
ALGOS = (A.A_NEIGHBORS, A.A_PRUNE, A.A_STRUCT)


def say(*a):
    print(*a, flush=True)


def build_both(ctx, pts, enc, ml):
    sx, keys, perm, c, r = O.sort_points(enc, pts)
    tree = O.Tree(keys, ml, enc, c, r)
    ctx.current = None
    ctx.set_points(pts)
    _, _, gperm, gkeys = ctx.sort_points(enc)
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)
    ctx.build(ml)
    return sx, tree


def check_all(ctx, sx, tree, centers, radii, kernels=KERNELS, algos=ALGOS, tag=""):
    """Lists (both precisions) and every count-mode variant against the oracle. Returns the FP32
    re-checks counted by the list kernels (the band was entered)."""
    rechecked = 0
    for kernel in kernels:
        for algo in algos:
            ref, sizes = O.radius(tree, sx, kernel, algo, centers, radii)
            for prec in (A.PREC_FP64, A.PREC_FP32):
                counts, lists, st = ctx.radius(kernel, algo, prec, centers=centers, radii=radii, lists=True)
                if prec == A.PREC_FP32:
                    rechecked += st["rechecked"]
                np.testing.assert_array_equal(counts, sizes, err_msg=f"{tag} lists k={kernel} a={algo} p={prec}")
                for q in range(len(centers)):
                    if not np.array_equal(lists[q], ref[q]):
                        raise AssertionError(f"{tag} k={kernel} a={algo} p={prec} query {q}: lists differ")
            try:
                for remap in (None, False):
                    ctx.set_remap(remap)
                    for g in (0, 16, 32, 64):
                        ctx.set_group(g)
                        for prec in (A.PREC_FP64, A.PREC_FP32):
                            counts, _ = ctx.radius(kernel, algo, prec, centers=centers, radii=radii)
                            np.testing.assert_array_equal(
                                counts, sizes, err_msg=f"{tag} counts k={kernel} a={algo} G={g} remap={remap} p={prec}")
            finally:
                ctx.set_group(-1)
                ctx.set_remap(None)
    return rechecked


@pytest.mark.gpu
@pytest.mark.parametrize("r", sorted(EXACT))
def test_band_points_on_the_radius(ctx, r):
    say(f"[gpu] band: points on the radius r={r}")
    pts, Q = boundary_cloud(r, seed=int(10 * r))
    sx, tree = build_both(ctx, pts, A.ENC_HILBERT, 16)
    radii = np.full(len(Q), r)
    rechecked = check_all(ctx, sx, tree, Q, radii, tag=f"r={r}")
    assert rechecked > 0, "no point fell in the FP32 band: the construction does not reach it"


@pytest.mark.gpu
@pytest.mark.parametrize("mag", [1e12, 3e14, 7e14, 1e15, 1.5e15, 1e16, 1e18, 1e19, 2e19, 6e19])
def test_band_magnitudes_near_the_overflow_guard(ctx, mag):
    say(f"[gpu] band: magnitude {mag:g}")
    pts, Q = magnitude_cloud(mag, seed=int(np.log10(mag) * 10))
    sx, tree = build_both(ctx, pts, A.ENC_MORTON, 32)
    r = float(np.float32(mag))  # the reference's radius is a float (main_options.hpp:177)
    check_all(ctx, sx, tree, Q, np.full(len(Q), r), kernels=(A.K_SPHERE, A.K_CIRCLE), tag=f"mag={mag:g}")


@pytest.mark.gpu
def test_band_node_judgement_equalities(ctx):
    say("[gpu] band: node judgement equalities")
    pts = node_boundary_cloud()
    sx, tree = build_both(ctx, pts, A.ENC_MORTON, 8)
    rl = node_radii(tree)
    Q, radii = node_boundary_queries(tree.centers, rl, count=400)
    check_all(ctx, sx, tree, Q, radii, tag="nodes")


def node_radii(tree):
    """Per-node radius (equal on every axis for the [0, 1024]^3 box): the level of node u follows
    from its prefix's placeholder bit (3 bits per level)."""
    pre = tree.prefixes
    level = np.zeros(len(pre), np.int64)
    for L in range(1, 22):  # the placeholder bit of a level-L prefix is bit 3 L
        level[pre >= np.uint64(1 << (3 * L))] = L
    return tree.radii[level, 0]


@pytest.mark.gpu
def test_band_knn_ties_at_the_kth_distance(ctx):
    say("[gpu] band: kNN ties")
    pts, Q = lattice_cloud()
    sx, tree = build_both(ctx, pts, A.ENC_HILBERT, 32)
    for k in (1, 2, 6, 7, 8, 9, 18, 19, 20, 26, 27, 32, 33, 57, 100, 123, 250):
        bi, bd = O.knn_brute(sx, Q, k)
        try:
            for lanes in [None] + [l for l in (2, 4, 8, 16, 32, 64) if ok_lanes(k, l)]:
                ctx.set_knn_lanes(lanes)
                for prec in (A.PREC_FP64, A.PREC_FP32):
                    gi, gd, _ = ctx.knn(k, precision=prec, centers=Q)
                    np.testing.assert_array_equal(gi.astype(np.int64), bi, err_msg=f"k={k} lanes={lanes} prec={prec}")
                    assert np.array_equal(gd.view(np.uint64), bd.view(np.uint64))
        finally:
            ctx.set_knn_lanes(None)


def ok_lanes(k, lanes):
    kc = 16
    while kc < k:
        kc *= 2
    offered = {16: (2, 4, 8), 32: (2, 4, 8), 64: (4, 8, 16), 128: (8, 16, 32), 256: (16, 32, 64)}.get(kc, (64,))
    return lanes in offered


# ------------------------------------------------------------------ CPU: the constructions hit their thresholds
def test_boundary_cloud_hits_the_radius():
    for r in EXACT:
        pts, Q = boundary_cloud(r, seed=int(10 * r), background=0)
        R = r * r
        on_s = on_c = 0
        for q in Q:
            near = pts[np.abs(pts - q).max(axis=1) <= 2 * r]
            s, c = d2_sphere(near, q), d2_circle(near, q)
            on_s += int(np.count_nonzero(s == R))
            on_c += int(np.count_nonzero(c == R))
            # the bisection pairs straddle the sphere's boundary: both sides are present
            assert np.any((s < R) & (s > R * (1 - 1e-12))) and np.any((s >= R) & (s < R * (1 + 1e-12)))
        assert on_s >= 4 * len(Q) and on_c >= 2 * len(Q)
        # circle `<=` and sphere `<` disagree on the exact points: the oracle sees it
        sx, keys, perm, c0, r0 = O.sort_points(A.ENC_HILBERT, pts)
        tree = O.Tree(keys, 16, A.ENC_HILBERT, c0, r0)
        flat = Q.copy()
        circle = O.radius_count(tree, sx, A.K_CIRCLE, A.A_PRUNE, flat, r)
        brute_lt = np.array([np.count_nonzero(d2_circle(sx, q) < R) for q in Q])
        assert np.all(circle >= brute_lt) and np.any(circle > brute_lt)


def test_node_boundary_queries_hit_equalities():
    pts = node_boundary_cloud()
    sx, keys, perm, c, r = O.sort_points(A.ENC_MORTON, pts)
    assert np.array_equal(c, [512.0, 512.0, 512.0]) and np.array_equal(r, [512.0, 512.0, 512.0])
    tree = O.Tree(keys, 8, A.ENC_MORTON, c, r)
    rl = node_radii(tree)
    assert set(np.unique(rl)).issubset({2.0 ** (9 - L) for L in range(22)})
    Q, radii = node_boundary_queries(tree.centers, rl, count=400)
    hits = 0
    for q, rr in zip(Q, radii):
        d = np.abs(q - tree.centers)
        e = d + rl[:, None]
        f = np.maximum(d - rl[:, None], 0.0)
        R = rr * rr
        hits += int(np.any(d == rr + rl[:, None]) or np.any(((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]) == R)
                    or np.any(((f[:, 0] * f[:, 0] + f[:, 1] * f[:, 1]) + f[:, 2] * f[:, 2]) == R))
    assert hits == len(Q)


def test_lattice_queries_tie_at_the_kth_distance():
    pts, Q = lattice_cloud(side=12)
    bi, bd = O.knn_brute(pts, Q[:50], 19)
    # the 19th neighbour of a lattice point (1 + 6 + 12 = 19) ends the sqrt(2) shell; k = 8 cuts it
    ties = [np.count_nonzero(bd[i] == bd[i, 7]) for i in range(len(bd))]
    assert max(ties) > 1
