"""The batch search pass on the device (oct_batch_queries + oct_batch_pipeline): the device build,
the grouping of the queries on the device and the work-queue count and kNN kernels, captured in one
HIP graph. Its results equal the per-call batch searches over the same forest (themselves checked
against the oracle) and a direct oracle sample; graph replays, the uncaptured sequence, replaced
points and autotuned kNN shapes keep them."""
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


def query_set(clouds, seed=3, nr=20000, nk=4000):
    rng = np.random.default_rng(seed)
    live = np.array([i for i, c in enumerate(clouds) if len(c[0])])
    rc = rng.choice(live, nr).astype(np.uint32)
    ri = np.array([rng.integers(0, len(clouds[c][0])) for c in rc], np.uint32)
    rr = rng.uniform(0.5, 5.0, nr).astype(np.float32).astype(np.float64)
    kc = rng.choice(live, nk).astype(np.uint32)
    ki = np.array([rng.integers(0, len(clouds[c][0])) for c in kc], np.uint32)
    kk = rng.integers(1, 300, nk).astype(np.uint32)
    kk[:40] = [1, 16, 17, 32, 33, 64, 65, 128, 129, 256, 257, 300] * 3 + [700, 1024, 1000, 512]
    return rc, ri, rr, kc, ki, kk


@pytest.fixture(scope="module")
def pipe():
    clouds = make_clouds(seed=44, count=10)
    clouds.append((clustered_duplicates(clusters=60, per=20, seed=8), A.ENC_MORTON, 16))
    b = A.Batch(0)
    for xyz, enc, ml in clouds:
        b.add(xyz, enc, ml)
    q = query_set(clouds)
    b.queries(*q, kmax=1024)
    say(f"[gpu] pipeline batch: {len(clouds)} clouds, {len(q[0])} radius + {len(q[3])} kNN queries")
    yield b, clouds, q
    b.close()


def per_call(b, q, kernel, algo, prec):
    rc, ri, rr, kc, ki, kk = q
    counts, _ = b.radius(kernel, algo, prec, rc, ri, rr)
    gi, gd, _ = b.knn(prec, kc, ki, kk, 1024)
    return counts, gi, gd


def same(a, b_):
    assert np.array_equal(a[0], b_[0]), "counts differ"
    assert np.array_equal(a[1], b_[1]), "kNN indices differ"
    assert np.array_equal(np.nan_to_num(a[2], nan=-1.0).view(np.uint64), np.nan_to_num(b_[2], nan=-1.0).view(np.uint64))


def test_pipeline_matches_per_call_and_oracle(pipe):
    b, clouds, q = pipe
    ms = b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
    say(f"[gpu] pipeline (graph) {ms:.2f} ms")
    res = b.pipeline_results()
    same(res, per_call(b, q, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32))
    rc, ri, rr, kc, ki, kk = q
    for c in (0, 3, 6, 10):
        sx, keys, perm, cc, r = O.sort_points(clouds[c][1], clouds[c][0])
        tree = O.Tree(keys, clouds[c][2], clouds[c][1], cc, r)
        sel = np.nonzero(rc == c)[0][:50]
        ref = np.array([O.radius_count(tree, sx, A.K_SPHERE, A.A_STRUCT, sx[[int(ri[j])]], float(rr[j]))[0] for j in sel])
        np.testing.assert_array_equal(res[0][sel], ref, err_msg=f"cloud {c}")
        ksel = np.nonzero(kc == c)[0][:30]
        if len(ksel):
            bi, bd = O.knn_brute(sx, sx[ki[ksel].astype(np.int64)], 1024)
            for row, j in enumerate(ksel):
                m = min(int(kk[j]), len(sx))
                np.testing.assert_array_equal(res[1][j, :m].astype(np.int64), bi[row, :m], err_msg=f"cloud {c} query {j}")
                assert np.all(res[1][j, m:] == -1)


def test_pipeline_replays_and_uncaptured(pipe):
    b, clouds, q = pipe
    b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
    ref = b.pipeline_results()
    for graph in (True, True, False):
        say(f"[gpu] pipeline graph={graph}")
        b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=graph)
        same(b.pipeline_results(), ref)


@pytest.mark.parametrize("kernel,algo,prec", [(A.K_SPHERE, A.A_PRUNE, A.PREC_FP64), (A.K_CUBE, A.A_NEIGHBORS, A.PREC_FP32),
                                              (A.K_CIRCLE, A.A_STRUCT, A.PREC_FP32)])
def test_pipeline_other_searches(pipe, kernel, algo, prec):
    b, clouds, q = pipe
    say(f"[gpu] pipeline kernel {kernel} algo {algo} prec {prec}")
    b.pipeline(kernel, algo, prec, graph=True)
    same(b.pipeline_results(), per_call(b, q, kernel, algo, prec))


def test_pipeline_autotuned_shapes(pipe):
    b, clouds, q = pipe
    b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
    ref = b.pipeline_results()
    lanes, ms = b.autotune_knn(A.PREC_FP32)
    say(f"[gpu] autotuned lanes {lanes}")
    b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
    same(b.pipeline_results(), ref)


def test_pipeline_new_points(pipe):
    b, clouds, q = pipe
    i = 3
    xyz = clouds[i][0]
    rng = np.random.default_rng(5)
    b.set_points(i, xyz + rng.normal(0.0, 0.2, xyz.shape))
    say("[gpu] pipeline replay with new points")
    b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
    same(b.pipeline_results(), per_call(b, q, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32))
    b.set_points(i, xyz)
    b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)


def test_pipeline_one_query_kind():
    """Only radius queries, then only kNN queries: the empty kind's work queues end at once."""
    clouds = make_clouds(seed=45, count=5)
    b = A.Batch(0)
    try:
        for xyz, enc, ml in clouds:
            b.add(xyz, enc, ml)
        rc, ri, rr, kc, ki, kk = query_set(clouds, seed=9, nr=3000, nk=800)
        b.queries(rc, ri, rr, [], [], [], kmax=0)
        b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, graph=True)
        counts, _, _ = b.pipeline_results()
        ref, _ = b.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, rc, ri, rr)
        np.testing.assert_array_equal(counts, ref)
        b.queries([], [], [], kc, ki, kk, kmax=1024)
        b.pipeline(A.K_SPHERE, A.A_STRUCT, A.PREC_FP64, graph=True)
        _, idx, d2 = b.pipeline_results()
        gi, gd, _ = b.knn(A.PREC_FP64, kc, ki, kk, 1024)
        np.testing.assert_array_equal(idx, gi)
    finally:
        b.close()
