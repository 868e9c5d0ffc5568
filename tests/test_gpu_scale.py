"""Mobile-laser-scan sweep (the generator of the 400 M-point stress workload, here at 20 M points):
Morton and Hilbert trees over the same cloud, all-points r = 0.5 m sphere counts in both precisions,
per-point equality across the two encoders (their trees are the same cells), and exact neighbour
sets on a crop of the cloud checked by brute force (a query farther than r from the crop's border
has the same neighbours in the crop as in the whole cloud)."""
import sys

import numpy as np
import pytest

import octree_amd as A

pytestmark = pytest.mark.gpu

N = 20_000_000
R = 0.5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mls():
    say(f"[gpu] mls sweep {N} points")
    return A.mls_sweep(N)


@pytest.fixture(scope="module")
def sweep_runs(ctx, mls):
    ctx.current = None
    ctx.set_points(mls)
    out = {}
    for enc in (A.ENC_MORTON, A.ENC_HILBERT):
        say(f"[gpu] mls encoder {enc}")
        c, r, perm, keys = ctx.sort_points(enc)
        info = ctx.build(128)
        c32, st32 = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, radius=R, nq=N)
        c64, st64 = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP64, radius=R, nq=N)
        out[enc] = dict(perm=perm, keys=keys, info=info, c32=c32, c64=c64, st32=st32, st64=st64)
    return out


def test_mls_keys_sorted_and_perm_bijective(sweep_runs):
    for enc, run in sweep_runs.items():
        assert np.all(run["keys"][1:] >= run["keys"][:-1]), f"encoder {enc}: keys not sorted"
        seen = np.zeros(N, np.bool_)
        seen[run["perm"].astype(np.int64)] = True
        assert seen.all(), f"encoder {enc}: permutation is not a bijection"


def test_mls_counts_fp32_eq_fp64_and_across_encoders(sweep_runs):
    per_point = {}
    for enc, run in sweep_runs.items():
        np.testing.assert_array_equal(run["c32"], run["c64"], err_msg=f"encoder {enc}")
        assert run["st32"]["total_neighbors"] == int(run["c64"].astype(np.int64).sum())
        orig = np.empty(N, np.uint32)
        orig[run["perm"].astype(np.int64)] = run["c32"]
        per_point[enc] = orig
    m, h = sweep_runs[A.ENC_MORTON], sweep_runs[A.ENC_HILBERT]
    for key in ("n_leaf", "n_internal", "n_total", "max_depth_seen"):
        assert m["info"][key] == h["info"][key], key
    np.testing.assert_array_equal(per_point[A.ENC_MORTON], per_point[A.ENC_HILBERT])
    say(f"[gpu] mls average neighbours {per_point[A.ENC_HILBERT].astype(np.int64).mean():.1f}")


def test_mls_crop_exact_sets(ctx, mls, sweep_runs):
    """neighbors (boxOverlap + isInside per point) lists on a 6 m x 6 m crop against brute force."""
    run = sweep_runs[A.ENC_HILBERT]  # the context holds the Hilbert tree (built last)
    sx = mls[run["perm"].astype(np.int64)]
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(6):
        a = sx[int(rng.integers(0, N))]
        x0, y0 = a[0] - 3.0, a[1] - 3.0
        box = (sx[:, 0] >= x0) & (sx[:, 0] <= x0 + 6.0) & (sx[:, 1] >= y0) & (sx[:, 1] <= y0 + 6.0)
        cand = np.nonzero(box)[0]
        inner = cand[(sx[cand, 0] > x0 + R + 0.05) & (sx[cand, 0] < x0 + 6.0 - R - 0.05) &
                     (sx[cand, 1] > y0 + R + 0.05) & (sx[cand, 1] < y0 + 6.0 - R - 0.05)]
        if len(inner) == 0:
            continue
        q = inner[rng.choice(len(inner), size=min(60, len(inner)), replace=False)]
        counts, lists, _ = ctx.radius(A.K_SPHERE, A.A_NEIGHBORS, A.PREC_FP32, idx=q.astype(np.uint32), radius=R,
                                      lists=True)
        P = sx[cand]
        for i, qi in enumerate(q):
            c = sx[qi]
            d1, d2, d3 = P[:, 0] - c[0], P[:, 1] - c[1], P[:, 2] - c[2]
            ref = cand[(d1 * d1 + d2 * d2) + d3 * d3 < R * R]
            np.testing.assert_array_equal(np.sort(lists[i].astype(np.int64)), ref, err_msg=f"query {qi}")
            assert counts[i] == len(ref)
            checked += 1
    assert checked >= 100
