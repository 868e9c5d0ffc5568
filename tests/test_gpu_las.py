"""LAS ingest on the device (oct_las_load): every LAS version and point format, extra bytes (records
staged in LDS and records read in place) and VLRs; coordinates and per-point attributes bit-exact with
the oracle's decode of the same file. A larger file is then sorted, built and searched exactly as
the same cloud handed over by oct_set_points."""
import sys

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O
from las_files import write_las

pytestmark = pytest.mark.gpu

CASES = [(minor, fmt) for minor in (2, 3) for fmt in range(6)] + [(4, fmt) for fmt in range(11)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("minor,fmt", CASES)
def test_las_load_matches_oracle(ctx, tmp_path, minor, fmt):
    rng = np.random.default_rng(7 * minor + fmt)
    n = 3000 + 97 * fmt  # partial last block
    X = rng.integers(-2 ** 31, 2 ** 31, n)
    Y = rng.integers(-2 ** 31, 2 ** 31, n)
    Z = rng.integers(-10 ** 6, 10 ** 6, n)
    extra = (0, 3, 70, 100)[fmt % 4]  # 100 extra bytes: records above 128 bytes, read in place
    path = str(tmp_path / f"v{minor}f{fmt}.las")
    write_las(path, X, Y, Z, fmt, minor=minor, scale=(0.001, 0.01, 0.0001), offset=(652000.25, -4100000.5, 12.0),
              extra=extra, vlrs=fmt % 3, seed=fmt + 10)
    say(f"[gpu] LAS 1.{minor} format {fmt} extra {extra}")
    h = ctx.las_load(path)
    info, s, o = O.las_header(path)
    assert h["n_points"] == info[3] == n and h["point_format"] == fmt and h["record_length"] == info[2]
    assert h["scale"] == list(s) and h["offset"] == list(o)
    ref = O.las_read(path)
    xyz, meta = ctx.las_fetch()
    assert np.array_equal(xyz.view(np.uint64), ref["xyz"].view(np.uint64))
    f = ref["fields"]
    for col, name in enumerate(("return_number", "number_of_returns", "scan_direction", "edge_of_flight_line",
                                "classification")):
        np.testing.assert_array_equal(meta[name], f[:, col], err_msg=name)
    np.testing.assert_array_equal(meta["scan_angle_rank"].view(np.uint8), f[:, 5])
    np.testing.assert_array_equal(meta["user_data"], f[:, 6])
    np.testing.assert_array_equal(meta["intensity"], ref["intensity"])
    np.testing.assert_array_equal(meta["point_source_id"], ref["psid"])
    for c, name in enumerate("rgb"):
        np.testing.assert_array_equal(meta[name], ref["rgb"][:, c], err_msg=name)


def test_las_cloud_searched_like_set_points(ctx, tmp_path):
    n = 2_000_000
    ref_xyz = A.lidar_tile(n, seed=12, extent=200.0, offset=(512000.0, 4100000.0))
    X = np.round((ref_xyz[:, 0] - 512000.0) / 0.001).astype(np.int64)
    Y = np.round((ref_xyz[:, 1] - 4100000.0) / 0.001).astype(np.int64)
    Z = np.round(ref_xyz[:, 2] / 0.001).astype(np.int64)
    path = str(tmp_path / "tile.las")
    write_las(path, X, Y, Z, 3, minor=2, scale=(0.001, 0.001, 0.001), offset=(512000.0, 4100000.0, 0.0))
    say("[gpu] LAS tile 2M points: load, sort, build, search")
    ctx.las_load(path)
    xyz, _ = ctx.las_fetch(meta=False)
    c1, r1, perm1, keys1 = ctx.sort_points(A.ENC_HILBERT)
    info1 = ctx.build(32)
    cnt1, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, radius=2.5, nq=n)
    ctx.set_points(O.las_read(path)["xyz"])
    c2, r2, perm2, keys2 = ctx.sort_points(A.ENC_HILBERT)
    info2 = ctx.build(32)
    cnt2, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, radius=2.5, nq=n)
    ctx.current = None
    np.testing.assert_array_equal(keys1, keys2)
    np.testing.assert_array_equal(perm1, perm2)
    assert info1["n_total"] == info2["n_total"]
    np.testing.assert_array_equal(cnt1, cnt2)


def test_las_edge_files(ctx, tmp_path):
    empty = str(tmp_path / "empty.las")
    write_las(empty, [], [], [], 0, minor=2)
    h = ctx.las_load(empty)
    assert h["n_points"] == 0
    xyz, meta = ctx.las_fetch()
    assert xyz.shape == (0, 3)
    laz = str(tmp_path / "c.laz")
    write_las(laz, [1], [2], [3], 1, compressed=True)
    with pytest.raises(A.OctError):
        ctx.las_load(laz)
    ctx.current = None
