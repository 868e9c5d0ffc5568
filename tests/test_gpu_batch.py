"""Batches of independent clouds: clouds of uneven size (uniform boxes and LiDAR-style tiles with their
own bounding boxes, Morton and Hilbert, different leaf sizes) searched together with a radius per
query and a k per query, against the oracle run on each cloud on its own."""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from batch_clouds import make_clouds

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch():
    clouds = make_clouds()
    b = A.Batch(0)
    for xyz, enc, ml in clouds:
        b.add(xyz, enc, ml)
    say(f"[gpu] batch of {len(clouds)} clouds, {sum(len(c[0]) for c in clouds)} points")
    b.build()
    refs = []
    for xyz, enc, ml in clouds:
        sx, keys, perm, c, r = O.sort_points(enc, xyz)
        refs.append((sx, O.Tree(keys, ml, enc, c, r)))
    yield b, clouds, refs
    b.close()


def test_batch_trees_match_single_cloud(batch):
    b, clouds, refs = batch
    for i, (sx, tree) in enumerate(refs):
        n, info = b.info(i)
        assert n == len(sx)
        assert info["n_total"] == tree.nTotal and info["n_leaf"] == tree.nLeaf, f"cloud {i}"


@pytest.mark.parametrize("kernel,algo", [(A.K_SPHERE, A.A_STRUCT), (A.K_SPHERE, A.A_NEIGHBORS), (A.K_CIRCLE, A.A_PRUNE),
                                         (A.K_CUBE, A.A_STRUCT), (A.K_SQUARE, A.A_NEIGHBORS)])
def test_batch_radius_mixed_radii(batch, kernel, algo):
    b, clouds, refs = batch
    rng = np.random.default_rng(kernel * 10 + algo)
    nq = 4000
    cloud = rng.integers(0, len(clouds), nq).astype(np.uint32)
    idx = np.array([rng.integers(0, len(clouds[c][0])) for c in cloud], np.uint32)
    radii = np.round(rng.uniform(0.5, 5.0, nq), 3)
    for prec in (A.PREC_FP64, A.PREC_FP32):
        counts, st = b.radius(kernel, algo, prec, cloud, idx, radii)
        for c in range(len(clouds)):
            sel = np.nonzero(cloud == c)[0]
            if len(sel) == 0:
                continue
            sx, tree = refs[c]
            ref = np.array([O.radius_count(tree, sx, kernel, algo, sx[[int(idx[q])]], float(radii[q]))[0] for q in sel])
            np.testing.assert_array_equal(counts[sel], ref, err_msg=f"cloud {c} prec {prec}")


def test_batch_knn_mixed_k(batch):
    b, clouds, refs = batch
    rng = np.random.default_rng(5)
    nq = 1500
    cloud = rng.integers(0, len(clouds), nq).astype(np.uint32)
    idx = np.array([rng.integers(0, len(clouds[c][0])) for c in cloud], np.uint32)
    k = rng.integers(1, 300, nq).astype(np.uint32)
    k[:20] = 1024
    kmax = int(k.max())
    for prec in (A.PREC_FP64, A.PREC_FP32):
        gi, gd, st = b.knn(prec, cloud, idx, k, kmax)
        for c in range(len(clouds)):
            sel = np.nonzero(cloud == c)[0]
            if len(sel) == 0:
                continue
            sx, tree = refs[c]
            kk = int(k[sel].max())
            bi, bd = O.knn_brute(sx, sx[idx[sel].astype(np.int64)], kk)
            for row, q in enumerate(sel):
                kq = int(k[q])
                m = min(kq, len(sx))
                np.testing.assert_array_equal(gi[q, :m].astype(np.int64), bi[row, :m], err_msg=f"cloud {c} query {q}")
                assert np.array_equal(gd[q, :m].view(np.uint64), bd[row, :m].view(np.uint64))
                assert np.all(gi[q, m:] == -1)
