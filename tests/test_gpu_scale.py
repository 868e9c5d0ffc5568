"""The mobile-laser-scan sweep at its stated size: 400 M points along a 5 km trajectory. Morton and
Hilbert trees over the same cloud, all-points r = 0.5 m sphere counts in both precisions, per-point
equality across the two encoders (their trees are the same cells), sorted keys and a bijective
permutation, and exact neighbour sets on crops of the cloud checked by brute force (a query farther
than r from the crop's border has the same neighbours in the crop as in the whole cloud)."""
import sys

import numpy as np
import pytest

import octree_amd as A

pytestmark = pytest.mark.gpu

N = 400_000_000
R = 0.5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sweep():
    say(f"[gpu] mls sweep {N} points")
    mls = A.mls_sweep(N)
    ctx = A.Octree(0)
    ctx.set_points(mls)
    runs = {}
    for enc in (A.ENC_MORTON, A.ENC_HILBERT):
        say(f"[gpu] mls encoder {enc}")
        c, r, perm, keys = ctx.sort_points(enc)
        info = ctx.build(128)
        c32, st32 = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, radius=R, nq=N)
        c64, st64 = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP64, radius=R, nq=N)
        say(f"[gpu] mls encoder {enc}: FP32 {st32['kernel_ms']:.0f} ms, FP64 {st64['kernel_ms']:.0f} ms")
        runs[enc] = dict(perm=perm, keys_sorted=bool(np.all(keys[1:] >= keys[:-1])), info=info, c32=c32, c64=c64,
                         st32=st32, st64=st64)
        del keys
    yield ctx, mls, runs  # the context holds the Hilbert tree (built last)
    ctx.close()


def test_mls_keys_sorted_and_perm_bijective(sweep):
    ctx, mls, runs = sweep
    for enc, run in runs.items():
        assert run["keys_sorted"], f"encoder {enc}: keys not sorted"
        seen = np.zeros(N, np.bool_)
        seen[run["perm"]] = True
        assert seen.all(), f"encoder {enc}: permutation is not a bijection"


def test_mls_counts_fp32_eq_fp64_and_across_encoders(sweep):
    ctx, mls, runs = sweep
    per_point = {}
    for enc, run in runs.items():
        assert np.array_equal(run["c32"], run["c64"]), f"encoder {enc}: FP32 + recheck differs from FP64"
        assert run["st32"]["total_neighbors"] == int(run["c64"].astype(np.int64).sum())
        orig = np.empty(N, np.uint32)
        orig[run["perm"]] = run["c32"]
        per_point[enc] = orig
    m, h = runs[A.ENC_MORTON], runs[A.ENC_HILBERT]
    for key in ("n_leaf", "n_internal", "n_total", "max_depth_seen"):
        assert m["info"][key] == h["info"][key], key
    assert np.array_equal(per_point[A.ENC_MORTON], per_point[A.ENC_HILBERT])
    say(f"[gpu] mls average neighbours {per_point[A.ENC_HILBERT].astype(np.int64).mean():.1f}")


def test_mls_crop_exact_sets(sweep):
    """neighbors (boxOverlap + isInside per point) lists on 6 m x 6 m crops against brute force."""
    ctx, mls, runs = sweep
    perm = runs[A.ENC_HILBERT]["perm"]
    inv = np.empty(N, np.uint32)  # input index -> position in the sorted cloud
    inv[perm] = np.arange(N, dtype=np.uint32)
    xs = np.ascontiguousarray(mls[:, 0])
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(4):
        a = mls[int(rng.integers(0, N))]
        x0, y0 = a[0] - 3.0, a[1] - 3.0
        sel = np.nonzero((xs >= x0) & (xs <= x0 + 6.0))[0]
        sel = sel[(mls[sel, 1] >= y0) & (mls[sel, 1] <= y0 + 6.0)]
        cand = np.sort(inv[sel].astype(np.int64))  # sorted-cloud positions of the crop's points
        P = mls[perm[cand].astype(np.int64)]
        inner = cand[(P[:, 0] > x0 + R + 0.05) & (P[:, 0] < x0 + 6.0 - R - 0.05) &
                     (P[:, 1] > y0 + R + 0.05) & (P[:, 1] < y0 + 6.0 - R - 0.05)]
        if len(inner) == 0:
            continue
        q = inner[rng.choice(len(inner), size=min(60, len(inner)), replace=False)]
        counts, lists, _ = ctx.radius(A.K_SPHERE, A.A_NEIGHBORS, A.PREC_FP32, idx=q.astype(np.uint32), radius=R,
                                      lists=True)
        for i, qi in enumerate(q):
            c = mls[int(perm[qi])]
            d1, d2, d3 = P[:, 0] - c[0], P[:, 1] - c[1], P[:, 2] - c[2]
            ref = cand[(d1 * d1 + d2 * d2) + d3 * d3 < R * R]
            np.testing.assert_array_equal(np.sort(lists[i].astype(np.int64)), ref, err_msg=f"query {qi}")
            assert counts[i] == len(ref)
            checked += 1
    assert checked >= 100
