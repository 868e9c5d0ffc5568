"""The allocation-free radius-lists entry point (oct_radius_lists_noalloc) against the allocating one
(oct_radius with lists + oct_radius_fetch), through the C ABI with caller-allocated device buffers:
two calls in a row (the second with fewer queries), all-points queries, and the refusals (an entries
buffer too small, too many queries)."""
import ctypes as C

import numpy as np
import pytest

import octree_amd as A

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers from the HIP runtime the library itself loaded."""

    def __init__(self):
        A.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(nbytes), 16)) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        if n:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def noalloc(ctx, hip, kernel, algo, prec, idx, radius, nq, capacity):
    d_idx = hip.put(idx) if idx is not None else None
    d_counts, d_offs = hip.alloc(4 * nq), hip.alloc(8 * (nq + 1))
    d_ent = hip.alloc(4 * max(capacity, 1))
    total = C.c_uint64(0)
    rc = ctx.L.oct_radius_lists_noalloc(ctx.h, kernel, algo, prec, d_idx, float(radius), nq, d_counts, d_offs, d_ent,
                                        capacity, C.byref(total))
    counts = hip.get(d_counts, nq, np.uint32)
    offs = hip.get(d_offs, nq + 1, np.uint64)
    words = int(total.value) * (2 if algo == A.A_STRUCT else 1)
    ent = hip.get(d_ent, min(words, capacity), np.uint32)
    return rc, counts, offs, ent, int(total.value)


def reference(ctx, kernel, algo, prec, idx, radius, nq):
    if idx is None:
        counts, lists, _ = ctx.radius(kernel, algo, prec, radius=radius, nq=nq, lists=True)
    else:
        counts, lists, _ = ctx.radius(kernel, algo, prec, idx=idx, radius=radius, lists=True)
    sizes = np.array([len(x) for x in lists], np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    ent = np.concatenate([x.reshape(-1) for x in lists] + [np.zeros(0, np.int64)]).astype(np.uint32)
    return counts, offs, ent


@pytest.fixture(scope="module")
def tree():
    xyz = A.lidar_tile(400_000, seed=21, extent=90.0)
    ctx = A.Octree(0)
    ctx.set_points(xyz)
    ctx.sort_points(A.ENC_HILBERT)
    ctx.build(32)
    hip = Hip()
    yield ctx, hip
    hip.free()
    ctx.close()


CASES = [(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, 2.5), (A.K_CUBE, A.A_PRUNE, A.PREC_FP64, 1.5),
         (A.K_CIRCLE, A.A_NEIGHBORS, A.PREC_FP32, 2.0), (A.K_SQUARE, A.A_STRUCT, A.PREC_FP64, 3.0)]


@pytest.mark.parametrize("kernel,algo,prec,radius", CASES)
def test_two_calls_match_the_allocating_entry(tree, kernel, algo, prec, radius):
    ctx, hip = tree
    rng = np.random.default_rng(kernel * 10 + algo)
    for nq in (300_000, 40_000):  # the second call smaller: the scratch keeps the first call's partials
        idx = rng.integers(0, ctx.n, nq).astype(np.uint32)
        print(f"[gpu] noalloc kernel={kernel} algo={algo} prec={prec} nq={nq}", flush=True)
        rcounts, roffs, rent = reference(ctx, kernel, algo, prec, idx, radius, nq)
        words = int(roffs[-1]) * (2 if algo == A.A_STRUCT else 1)
        rc, counts, offs, ent, total = noalloc(ctx, hip, kernel, algo, prec, idx, radius, nq, words)
        hip.free()
        assert rc == 0, A.lib().oct_last_error(ctx.h)
        np.testing.assert_array_equal(counts, rcounts)
        np.testing.assert_array_equal(offs, roffs)
        assert total == int(roffs[-1])
        np.testing.assert_array_equal(ent, rent)


def test_all_points_and_refusals(tree):
    ctx, hip = tree
    nq = ctx.n
    print(f"[gpu] noalloc all points nq={nq}", flush=True)
    rcounts, roffs, rent = reference(ctx, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, None, 2.5, nq)
    words = 2 * int(roffs[-1])
    rc, counts, offs, ent, total = noalloc(ctx, hip, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, None, 2.5, nq, words)
    hip.free()
    assert rc == 0
    np.testing.assert_array_equal(counts, rcounts)
    np.testing.assert_array_equal(offs, roffs)
    np.testing.assert_array_equal(ent, rent)
    # an entries buffer one word short: refused after counts, offsets and the total are written
    rc, counts, offs, _, total = noalloc(ctx, hip, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, None, 2.5, nq, words - 1)
    hip.free()
    assert rc != 0
    assert b"entries_capacity" in A.lib().oct_last_error(ctx.h)
    np.testing.assert_array_equal(counts, rcounts)
    np.testing.assert_array_equal(offs, roffs)
    assert total == int(roffs[-1])
    # more queries than the scan scratch holds: refused before any launch
    d = hip.alloc(16)
    t = C.c_uint64(0)
    rc = ctx.L.oct_radius_lists_noalloc(ctx.h, A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, d, 2.5, (1 << 26) + 1, d, d, d, 0,
                                        C.byref(t))
    hip.free()
    assert rc != 0 and b"allocation-free" in A.lib().oct_last_error(ctx.h)
    # zero queries: offsets[0] = 0, nothing else written
    rc, counts, offs, ent, total = noalloc(ctx, hip, A.K_SPHERE, A.A_PRUNE, A.PREC_FP64, np.zeros(0, np.uint32), 2.5, 0, 0)
    hip.free()
    assert rc == 0 and total == 0 and offs.tolist() == [0]
