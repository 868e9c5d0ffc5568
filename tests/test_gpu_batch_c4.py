"""The batched workload at its stated size: 256 independent clouds (~200 M points, sizes log-uniform in
[1e4, 4e6], uniform boxes and airborne tiles at random origins), 50 k centres per cloud (10.24 M radius
queries, r uniform in [0.5, 5] m; 2.56 M kNN queries, k uniform in {8..256}), built and searched by the
device pipeline in one HIP graph (oct_batch_queries + oct_batch_pipeline).

Checked: FP32 + recheck equals FP64 for every query (counts, kNN indices and squared distances); on
sampled clouds (the smallest, the largest and two others) the cloud's sorted keys, permutation and
tree equal the oracle's run on that cloud alone, its radius counts equal the oracle's, its kNN squared
distances equal the oracle's best-first search, and its kNN index lists equal a brute force over a
crop on a few queries."""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from crop_knn import crop_brute_knn

pytestmark = pytest.mark.gpu

ML = 128


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4():
    W = A.BatchWorkload()
    order = np.argsort(W.sizes)
    sample = sorted({int(order[0]), int(order[-1]), 3, 128})
    b = A.Batch(0)
    kept = {}
    for i in range(W.count):
        xyz = W.cloud(i)
        b.add(xyz, A.ENC_HILBERT, ML)
        if i in sample:
            kept[i] = xyz
    rc, ri, rr = W.radius_queries()
    kc, ki, kk = W.knn_queries()
    b.queries(rc, ri, rr, kc, ki, kk, kmax=256)
    say(f"[gpu] C4 batch: {W.count} clouds, {W.points} points, {len(rc)} radius + {len(kc)} kNN queries")
    res = {}
    for name, prec in (("fp32", A.PREC_FP32), ("fp64", A.PREC_FP64)):
        ms = b.pipeline(A.K_SPHERE, A.A_STRUCT, prec, graph=True)
        say(f"[gpu] C4 pipeline {name}: {ms:.1f} ms (first replay, with capture)")
        res[name] = b.pipeline_results()
    yield b, W, kept, res
    b.close()


def test_c4_fp32_equals_fp64(c4):
    b, W, kept, res = c4
    (c32, i32, d32), (c64, i64, d64) = res["fp32"], res["fp64"]
    assert np.array_equal(c32, c64), f"counts differ on {int(np.count_nonzero(c32 != c64))} queries"
    assert np.array_equal(i32, i64), "kNN indices differ"
    assert np.array_equal(np.nan_to_num(d32, nan=-1.0).view(np.uint64), np.nan_to_num(d64, nan=-1.0).view(np.uint64))
    kc, ki, kk = W.knn_queries()
    # rows hold exactly min(n, k) neighbours, then -1 padding
    filled = (i64 >= 0).sum(axis=1)
    np.testing.assert_array_equal(filled, np.minimum(kk, W.sizes[kc]))


def test_c4_sampled_clouds_against_oracle(c4):
    b, W, kept, res = c4
    counts, kidx, kd2 = res["fp32"]
    rc, ri, rr = W.radius_queries()
    kc, ki, kk = W.knn_queries()
    for c, xyz in kept.items():
        say(f"[gpu] C4 cloud {c}: {len(xyz)} points against the oracle")
        sx, keys, perm, cc, r = O.sort_points(A.ENC_HILBERT, xyz)
        tree = O.Tree(keys, ML, A.ENC_HILBERT, cc, r)
        gkeys, gperm = b.cloud_sorted(c)
        np.testing.assert_array_equal(gkeys, keys)
        np.testing.assert_array_equal(gperm, perm)
        g = b.cloud_tree(c)
        np.testing.assert_array_equal(g["prefixes"], tree.prefixes)
        np.testing.assert_array_equal(g["ranges"], tree.ranges)
        assert np.array_equal(g["centers"].view(np.uint64), tree.centers.view(np.uint64))
        sel = np.nonzero(rc == c)[0][:400]
        ref = np.array([O.radius_count(tree, sx, A.K_SPHERE, A.A_STRUCT, sx[[int(ri[j])]], float(rr[j]))[0] for j in sel])
        np.testing.assert_array_equal(counts[sel], ref, err_msg=f"cloud {c}")
        ksel = np.nonzero(kc == c)[0][:200]
        for j in ksel:
            k = int(kk[j])
            oi, od = O.knn(tree, sx, sx[[int(ki[j])]], k)
            m = min(k, len(sx))
            assert np.array_equal(kd2[j, :m].view(np.uint64), od[0, :m].view(np.uint64)), f"cloud {c} query {j}"
        for j in ksel[:4]:
            k = min(int(kk[j]), len(sx))
            oi, od = O.knn(tree, sx, sx[[int(ki[j])]], k)
            bi, bd = crop_brute_knn(sx, tree, sx[int(ki[j])], k, float(od[0, k - 1]))
            np.testing.assert_array_equal(kidx[j, :k].astype(np.int64), bi, err_msg=f"cloud {c} query {j}")
