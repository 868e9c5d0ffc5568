"""The reorder of sortPoints applied to caller records (oct_gather_records: the Point container and the
PointMetadata vector moving with the keys, point_encoder.hpp:321-323, 367-375): records of 32 B
(16-byte words), 12 B (4-byte words) and 7 B (bytes), host buffers and device buffers, equal to the
oracle's permutation applied on the host."""
import numpy as np
import pytest

import octree_amd as A
import pyoracle as O

pytestmark = pytest.mark.gpu


def test_gather_records_host(ctx):
    xyz = A.uniform_box(200_003, seed=9)
    ctx.set_points(xyz)
    ctx.current = None
    _, _, perm, _ = ctx.sort_points(A.ENC_HILBERT)
    _, _, ref_perm, _, _ = O.sort_points(A.ENC_HILBERT, xyz)
    np.testing.assert_array_equal(perm, ref_perm)
    p = ref_perm.astype(np.int64)
    rng = np.random.default_rng(2)
    meta = np.zeros(len(xyz), A.POINT_META)
    meta["intensity"] = rng.uniform(0, 65535, len(xyz))
    meta["classification"] = rng.integers(0, 32, len(xyz))
    meta["r"] = rng.integers(0, 65536, len(xyz))
    for rec in (meta, rng.integers(0, 2 ** 31, (len(xyz), 3)).astype(np.int32),
                rng.integers(0, 256, (len(xyz), 7)).astype(np.uint8), xyz):
        got = ctx.gather_records(rec)
        assert np.array_equal(got.view(np.uint8), rec[p].view(np.uint8))


def test_gather_records_device(ctx):
    """Device buffers allocated with the HIP runtime itself (hipMalloc through ctypes)."""
    import ctypes as C
    A.lib()
    hip = C.CDLL("libamdhip64.so.7")  # the HIP runtime the library itself loaded
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    xyz = A.uniform_box(100_000, seed=10)
    ctx.set_points(xyz)
    ctx.current = None
    _, _, perm, _ = ctx.sort_points(A.ENC_MORTON)
    src = np.arange(len(xyz) * 8, dtype=np.int32).reshape(len(xyz), 8)
    dst = np.zeros_like(src)
    d_src, d_dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_src), src.nbytes) == 0 and hip.hipMalloc(C.byref(d_dst), src.nbytes) == 0
    try:
        assert hip.hipMemcpy(d_src, src.ctypes.data, src.nbytes, 1) == 0  # host to device
        ctx.gather_records_dev(d_src, d_dst, 32)
        assert hip.hipMemcpy(dst.ctypes.data, d_dst, dst.nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(d_src), hip.hipFree(d_dst)
    np.testing.assert_array_equal(dst, src[perm.astype(np.int64)])
