"""GPU sort parity: every sort path of oct_sort against the oracle's stable sort, on clouds built to
reach each branch of the partition sort.

The partition sort (OCT_TUNE_SORT 1 and 2, the default for more than 2048 / 4096 keys) orders the keys
by their top 16 bits in two device-wide passes, then sorts each 16-bit segment inside one workgroup,
splitting segments above the workgroup's LDS capacity by the next 8 bits until the pieces fit; a
segment above 32768 keys sends the whole sort to the LSD passes (OCT_TUNE_SORT 0). Bar: permutation
and sorted keys bit-exact (the stable sort by key, reference point_encoder.hpp:285-383).
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


def clustered(n_cluster, n_spread, width, seed):
    """n_cluster points inside a cube of side `width` at one spot, n_spread uniform over 1000 m."""
    rng = np.random.default_rng(seed)
    c = np.array([412.5, 377.25, 21.0]) + width * rng.random((n_cluster, 3))
    u = np.array([0.0, 0.0, -10.0]) + np.array([1000.0, 1000.0, 60.0]) * rng.random((n_spread, 3))
    xyz = np.concatenate([c, u])
    return xyz[rng.permutation(len(xyz))]


def duplicates(seed):
    """100 k points at 37 distinct spots, plus 20 k copies of one more: long runs of equal keys."""
    rng = np.random.default_rng(seed)
    spots = 500.0 * rng.random((38, 3))
    idx = np.concatenate([rng.integers(0, 37, 100_000), np.full(20_000, 37)])
    rng.shuffle(idx)
    return spots[idx]


def spots(seed):
    """Ten spots of 6-9 k points within 1 cm each, plus 100 k uniform points over 1000 m."""
    rng = np.random.default_rng(seed)
    parts = [np.array([0.0, 0.0, -10.0]) + np.array([1000.0, 1000.0, 60.0]) * rng.random((100_000, 3))]
    for _ in range(10):
        c = np.array([1000.0, 1000.0, 40.0]) * rng.random(3)
        parts.append(c + 0.01 * rng.random((int(rng.integers(6000, 9000)), 3)))
    xyz = np.concatenate(parts)
    return xyz[rng.permutation(len(xyz))]


CLOUDS = {
    # 16-bit segments of a few thousand keys: mostly runs of small segments, some split once
    "tile_1m_hilbert": lambda: (A.lidar_tile(1_000_000, seed=7, extent=300.0), A.ENC_HILBERT),
    "uniform_300k_morton": lambda: (A.uniform_box(300_000, seed=5), A.ENC_MORTON),
    # one 16-bit segment of 20 k / 30 k keys spread over a few metres: split inside one workgroup into
    # many pieces, sorted in runs
    "cluster_20k_hilbert": lambda: (clustered(20_000, 100_000, 2.0, 21), A.ENC_HILBERT),
    "cluster_30k_morton": lambda: (clustered(30_000, 50_000, 5.0, 22), A.ENC_MORTON),
    # several dense spots, each a segment of 6-9 k keys within a centimetre: splits down several levels
    "spots_hilbert": lambda: (spots(23), A.ENC_HILBERT),
    # a segment above 32768 keys: the LSD passes take over after the first partition pass
    "cluster_60k_hilbert": lambda: (clustered(60_000, 200_000, 0.01, 11), A.ENC_HILBERT),
    "cluster_150k_morton": lambda: (clustered(150_000, 50_000, 0.001, 12), A.ENC_MORTON),
    # runs of equal keys longer than an LDS sort: splits until the key bits run out
    "duplicates_120k_hilbert": lambda: (duplicates(13), A.ENC_HILBERT),
    # every key equal (zero-extent box)
    "same_30k_morton": lambda: (np.tile([[3.25, -7.5, 1.0]], (30_000, 1)), A.ENC_MORTON),
}
SIZES = [1, 2, 100, 2047, 2048, 2049, 4095, 4096, 4097, 8193, 12289, 70_001]


@pytest.fixture(scope="module")
def restore_sort(ctx):
    yield
    ctx.set_sort(None)
    ctx.current = None


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(CLOUDS))
def test_sort_paths_match_oracle(ctx, restore_sort, name, mode):
    xyz, enc = CLOUDS[name]()
    say(f"[sort] {name} n={len(xyz)} mode={mode}")
    _, keys, perm, c, r = O.sort_points(enc, xyz)
    ctx.set_sort(mode)
    ctx.set_points(xyz)
    ctx.current = None
    gc, gr, gperm, gkeys = ctx.sort_points(enc)
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n", SIZES)
def test_sort_sizes_around_capacity(ctx, restore_sort, n, mode):
    xyz = A.uniform_box(n, seed=100 + n % 97)
    say(f"[sort] uniform n={n} mode={mode}")
    _, keys, perm, c, r = O.sort_points(A.ENC_HILBERT, xyz)
    ctx.set_sort(mode)
    ctx.set_points(xyz)
    ctx.current = None
    gc, gr, gperm, gkeys = ctx.sort_points(A.ENC_HILBERT)
    np.testing.assert_array_equal(gkeys, keys)
    np.testing.assert_array_equal(gperm, perm)


def test_sort_knob_rejects_bad_values(ctx):
    with pytest.raises(A.OctError):
        ctx.set_sort(3)
    ctx.set_sort(None)


def plane_cloud(n, seed):
    """Points on the plane z = 0 (a zero-extent axis), negative and positive x and y."""
    rng = np.random.default_rng(seed)
    xyz = np.zeros((n, 3))
    xyz[:, :2] = 200.0 * rng.random((n, 2)) - 100.0
    return xyz


@pytest.mark.parametrize("name,make", [
    ("uniform_3", lambda: A.uniform_box(3, seed=5)),
    ("uniform_4097", lambda: A.uniform_box(4097, seed=6)),
    ("tile_300k", lambda: A.lidar_tile(300_000, seed=7, extent=120.0)),
    ("plane_50k", lambda: plane_cloud(50_000, 8)),
])
@pytest.mark.parametrize("mode", [0, 1])
def test_sort_points_equals_separate_calls(ctx, restore_sort, name, make, mode):
    """oct_sort_points keeps the box on the device between bbox, encode and sort; the separate calls
    hand it through the host. Same box (bit for bit), keys and permutation."""
    xyz = make()
    say(f"[sort] sort_points vs separate calls {name} mode={mode}")
    ctx.set_sort(mode)
    ctx.set_points(xyz)
    ctx.current = None
    c1, r1, p1, k1 = ctx.sort_points(A.ENC_HILBERT)
    ctx.set_points(xyz)
    c2, r2 = ctx.bbox()
    ctx.encode(A.ENC_HILBERT, c2, r2, fetch=False)
    p2, k2 = ctx.sort()
    assert np.array_equal(c1.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_array_equal(p1, p2)


def test_default_sort_remembers_fallback(ctx, restore_sort):
    """Default form on a cloud whose dense cluster makes a 16-bit segment above the workgroup limit:
    the first sort falls back to the LSD passes, the next sorts of the same points run them directly,
    and new points start from the partition sort again - every result the oracle's."""
    ctx.set_sort(None)
    for name in ("cluster_60k_hilbert", "uniform_300k_morton", "cluster_150k_morton"):
        xyz, enc = CLOUDS[name]()
        _, keys, perm, c, r = O.sort_points(enc, xyz)
        ctx.set_points(xyz)
        ctx.current = None
        for rep in range(2):
            say(f"[sort] default form {name} rep {rep}")
            gc, gr, gperm, gkeys = ctx.sort_points(enc)
            np.testing.assert_array_equal(gkeys, keys)
            np.testing.assert_array_equal(gperm, perm)
