"""Device-driven batch build (oct_batch_build_device, captured in a HIP graph): every cloud's sorted
keys, permutation and tree arrays against the oracle run on that cloud alone; graph replays (with
and without replaced points) and the uncaptured launch sequence give the same forest; a cloud of
clustered duplicates outgrows the first node-capacity estimate (the grow-and-rebuild path); searches
over the device-built forest equal those over the per-cloud-built forest and the oracle."""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from batch_clouds import clustered_duplicates, make_clouds

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_clouds():
    clouds = make_clouds(seed=33, count=12)
    clouds.append((clustered_duplicates(), A.ENC_HILBERT, 16))
    clouds.append((np.tile([[3.0, -2.0, 7.5]], (300, 1)), A.ENC_MORTON, 16))  # one point repeated: depth limit
    clouds.append((np.zeros((0, 3)), A.ENC_HILBERT, 32))  # empty cloud
    return clouds


def oracle_cloud(xyz, enc, ml):
    sx, keys, perm, c, r = O.sort_points(enc, xyz)
    return sx, keys, perm, O.Tree(keys, ml, enc, c, r)


@pytest.fixture(scope="module")
def dev():
    clouds = device_clouds()
    b = A.Batch(0)
    for xyz, enc, ml in clouds:
        b.add(xyz, enc, ml)
    say(f"[gpu] device batch of {len(clouds)} clouds, {sum(len(c[0]) for c in clouds)} points (graph build)")
    ms = b.build_device(graph=True)
    say(f"[gpu] device build {ms:.3f} ms")
    refs = [oracle_cloud(*c) if len(c[0]) else None for c in clouds]
    yield b, clouds, refs
    b.close()


def check_cloud(b, i, ref):
    sx, keys, perm, tree = ref
    n, info = b.info(i)
    assert n == len(sx)
    gk, gp = b.cloud_sorted(i)
    np.testing.assert_array_equal(gk, keys, err_msg=f"cloud {i} keys")
    np.testing.assert_array_equal(gp, perm, err_msg=f"cloud {i} permutation")
    assert info["n_leaf"] == tree.nLeaf and info["n_internal"] == tree.nInternal, f"cloud {i}"
    assert info["n_total"] == tree.nTotal and info["max_depth_seen"] == tree.maxDepthSeen, f"cloud {i}"
    g = b.cloud_tree(i)
    np.testing.assert_array_equal(g["prefixes"], tree.prefixes, err_msg=f"cloud {i}")
    np.testing.assert_array_equal(g["offsets"], tree.offsets[:tree.nTotal], err_msg=f"cloud {i}")
    np.testing.assert_array_equal(g["ranges"], tree.ranges, err_msg=f"cloud {i}")
    assert np.array_equal(g["centers"].view(np.uint64), tree.centers.view(np.uint64)), f"cloud {i}: centres"
    assert np.array_equal(g["radii"].view(np.uint64), tree.radii.view(np.uint64)), f"cloud {i}: radii"


def test_device_build_matches_oracle(dev):
    b, clouds, refs = dev
    for i, ref in enumerate(refs):
        if ref is None:
            n, info = b.info(i)
            assert n == 0 and info["n_total"] == 1 and info["n_leaf"] == 1 and info["max_depth_seen"] == 0
            continue
        check_cloud(b, i, ref)
    # the repeated point splits down to level 21: 21 internal nodes
    n, info = b.info(len(clouds) - 2)
    assert info["n_internal"] == 21 and info["max_depth_seen"] == 21


def test_replays_and_uncaptured_sequence_agree(dev):
    b, clouds, refs = dev
    first = [b.cloud_sorted(i) for i in range(len(clouds))]
    trees = [b.cloud_tree(i)["offsets"] for i in range(len(clouds))]
    for graph in (True, True, False):
        say(f"[gpu] device rebuild graph={graph}")
        b.build_device(graph=graph)
        for i in range(len(clouds)):
            k, p = b.cloud_sorted(i)
            np.testing.assert_array_equal(k, first[i][0])
            np.testing.assert_array_equal(p, first[i][1])
            np.testing.assert_array_equal(b.cloud_tree(i)["offsets"], trees[i])


def test_replay_after_new_points(dev):
    b, clouds, refs = dev
    i = 5
    xyz = clouds[i][0]
    rng = np.random.default_rng(77)
    moved = xyz[rng.permutation(len(xyz))] + rng.normal(0.0, 0.3, xyz.shape)
    b.set_points(i, moved)
    say("[gpu] replay with new points of one cloud")
    b.build_device(graph=True)
    check_cloud(b, i, oracle_cloud(moved, clouds[i][1], clouds[i][2]))
    check_cloud(b, 4, refs[4])
    b.set_points(i, xyz)
    b.build_device(graph=True)
    check_cloud(b, i, refs[i])


@pytest.fixture(scope="module")
def per_cloud(dev):
    _, clouds, _ = dev
    b = A.Batch(0)
    for xyz, enc, ml in clouds:
        b.add(xyz, enc, ml)
    say("[gpu] per-cloud batch of the same clouds")
    b.build()
    yield b
    b.close()


def test_device_forest_searches(dev, per_cloud):
    b, clouds, refs = dev
    rng = np.random.default_rng(9)
    live = [i for i, c in enumerate(clouds) if len(c[0])]
    nq = 6000
    cloud = rng.choice(live, nq).astype(np.uint32)
    idx = np.array([rng.integers(0, len(clouds[c][0])) for c in cloud], np.uint32)
    radii = np.round(rng.uniform(0.5, 5.0, nq), 3)
    for kernel, algo in ((A.K_SPHERE, A.A_STRUCT), (A.K_CUBE, A.A_NEIGHBORS)):
        for prec in (A.PREC_FP64, A.PREC_FP32):
            cd, _ = b.radius(kernel, algo, prec, cloud, idx, radii)
            cp, _ = per_cloud.radius(kernel, algo, prec, cloud, idx, radii)
            np.testing.assert_array_equal(cd, cp, err_msg=f"kernel {kernel} algo {algo} prec {prec}")
    for c in live[:6]:
        sel = np.nonzero(cloud == c)[0][:40]
        sx, _, _, tree = refs[c]
        ref = np.array([O.radius_count(tree, sx, A.K_SPHERE, A.A_STRUCT, sx[[int(idx[q])]], float(radii[q]))[0] for q in sel])
        cd, _ = b.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, cloud[sel], idx[sel], radii[sel])
        np.testing.assert_array_equal(cd, ref, err_msg=f"cloud {c}")
    k = rng.integers(1, 200, nq).astype(np.uint32)
    for prec in (A.PREC_FP64, A.PREC_FP32):
        gi, gd, _ = b.knn(prec, cloud, idx, k, 200)
        pi, pd, _ = per_cloud.knn(prec, cloud, idx, k, 200)
        np.testing.assert_array_equal(gi, pi)
        assert np.array_equal(np.nan_to_num(gd, nan=-1.0).view(np.uint64), np.nan_to_num(pd, nan=-1.0).view(np.uint64))


def test_capacity_grows_and_rebuilds():
    """Only clouds of clustered duplicates: the forest needs several times the first node-capacity
    estimate, so the build overflows, grows its node storage and runs again (graph re-captured)."""
    clouds = [(clustered_duplicates(clusters=150, per=17, seed=s), A.ENC_HILBERT if s % 2 else A.ENC_MORTON, 16)
              for s in (1, 2)]
    b = A.Batch(0)
    try:
        for xyz, enc, ml in clouds:
            b.add(xyz, enc, ml)
        say("[gpu] clustered-duplicates batch (capacity growth)")
        b.build_device(graph=True)
        for i, c in enumerate(clouds):
            ref = oracle_cloud(*c)
            assert ref[3].nTotal > 4 * (64 + 8 * len(c[0]) // 16)  # far beyond the estimate
            check_cloud(b, i, ref)
        b.build_device(graph=True)
        check_cloud(b, 1, oracle_cloud(*clouds[1]))
    finally:
        b.close()


def test_clouds_fixed_once_built_or_queried():
    """The device build and the query grouping size their buffers for the clouds present: adding a
    cloud afterwards is refused (also after oct_batch_set_points, which clears the built state)."""
    a, b_ = A.uniform_box(5000, seed=1), A.uniform_box(3000, seed=2)
    b = A.Batch(0)
    try:
        b.add(a, A.ENC_HILBERT, 32)
        b.build_device(graph=False)
        with pytest.raises(A.OctError):
            b.add(b_, A.ENC_HILBERT, 32)
        b.set_points(0, a[::-1])
        with pytest.raises(A.OctError):
            b.add(b_, A.ENC_HILBERT, 32)
    finally:
        b.close()
    b = A.Batch(0)
    try:
        b.add(a, A.ENC_HILBERT, 32)
        b.queries([0], [1], [1.0], [0], [2], [3], kmax=3)
        with pytest.raises(A.OctError):
            b.add(b_, A.ENC_HILBERT, 32)
        b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
        counts, idx, d2 = b.pipeline_results()
        assert counts[0] >= 1 and idx[0, 0] == 2
    finally:
        b.close()
