"""GPU tree build against the oracle's, array by array, on every path of the build.

The build runs its level loop on the device (a chained scan of the split flags and the children of each
level, the level table read back once) when the tree fits the node arrays sized for 8 n / maxLeaf + 64
nodes and is at most 12 levels deep; otherwise it redoes the build with the host level loop, which
grows the arrays. Bar: prefixes, offsets, ranges, FP64 centres and radii bit-exact (reference
linear_octree.hpp:151-565).
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


def check_build(ctx, xyz, enc, ml):
    _, keys, perm, c, r = O.sort_points(enc, xyz)
    tree = O.Tree(keys, ml, enc, c, r)
    ctx.set_points(xyz)
    ctx.current = None
    ctx.sort_points(enc)
    info = ctx.build(ml)
    g = ctx.get_tree()
    assert info["n_leaf"] == tree.nLeaf
    assert info["n_internal"] == tree.nInternal
    assert info["n_total"] == tree.nTotal
    assert info["max_depth_seen"] == tree.maxDepthSeen
    np.testing.assert_array_equal(g["prefixes"], tree.prefixes)
    np.testing.assert_array_equal(g["offsets"], tree.offsets[:tree.nTotal])
    np.testing.assert_array_equal(g["ranges"], tree.ranges)
    assert np.array_equal(g["centers"].view(np.uint64), tree.centers.view(np.uint64)), "centres not bit-exact"
    assert np.array_equal(g["radii"].view(np.uint64), tree.radii.view(np.uint64))
    return info


def close_pairs(n_pairs, gap, n_spread, seed):
    """n_pairs pairs of points `gap` apart, plus n_spread uniform points over 1000 m."""
    rng = np.random.default_rng(seed)
    a = 1000.0 * rng.random((n_pairs, 3))
    b = a + np.array([gap, 0.0, 0.0])
    u = 1000.0 * rng.random((n_spread, 3))
    xyz = np.concatenate([a, b, u])
    return xyz[rng.permutation(len(xyz))]


@pytest.mark.parametrize("ml", [16, 32, 128])
def test_build_device_levels(ctx, ml):
    """A tile whose tree fits the arrays and 12 levels: the device level loop."""
    say(f"[build] tile ml={ml}")
    info = check_build(ctx, A.lidar_tile(300_000, seed=9, extent=120.0), A.ENC_HILBERT, ml)
    assert info["max_depth_seen"] <= 12


def test_build_tile_single_point_leaves(ctx):
    """The same tile with maxLeaf 1: points repeated on the 1 mm grid take it to 17 levels (host loop)."""
    say("[build] tile ml=1")
    info = check_build(ctx, A.lidar_tile(300_000, seed=9, extent=120.0), A.ENC_HILBERT, 1)
    assert info["max_depth_seen"] > 12


def test_build_deep_tree_within_capacity(ctx):
    """One pair 1 cm apart among 2000 spread points, maxLeaf 1: 17 levels, the host loop takes over."""
    say("[build] deep pair")
    info = check_build(ctx, close_pairs(1, 0.01, 2000, 31), A.ENC_MORTON, 1)
    assert info["max_depth_seen"] > 12


def test_build_overflows_capacity(ctx):
    """50 pairs 1 um apart, maxLeaf 1: eight nodes per pair and level outgrow 8 n / maxLeaf + 64."""
    say("[build] overflow")
    info = check_build(ctx, close_pairs(50, 1e-6, 0, 32), A.ENC_HILBERT, 1)
    assert info["n_total"] > 64 + 8 * 100


def test_build_repeated_and_shrinking(ctx):
    """Builds in a row on one context, larger then smaller, each against the oracle."""
    for n, ml in ((200_000, 32), (5_000, 8), (120_000, 64)):
        say(f"[build] n={n} ml={ml}")
        check_build(ctx, A.uniform_box(n, seed=n), A.ENC_HILBERT, ml)
