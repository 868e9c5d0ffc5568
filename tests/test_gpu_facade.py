"""The drop-in C++ facade (octrees-benchmark_amd/facade/, reference-named headers) driven the way the
reference's example and tests drive LinearOctree (examples/example.cpp; tests/structures/*.cpp):
sortPoints, the LinearOctree constructor, the three searches over the four kernels (by kernel_t and by
kernel object), knn, the batch count, singleTraversal, the key codec, encodePoints / getCenter and the
PointsSoA container. Every printed result is compared with the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import octree_amd as A
import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "octrees-benchmark_amd", "facade_check")


def parse(text):
    out = {"pt": [], "struct": {}, "prune": {}, "neighbors": {}, "knn": {}, "approx": {}}
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "box":
            out["box"] = np.array([float(v) for v in f[1:]])
        elif f[0] == "pt":
            out["pt"].append((int(f[2]), int(f[3])))
        elif f[0] == "tree":
            out["tree"] = tuple(int(v) for v in f[1:])
        elif f[0] in ("struct", "prune", "neighbors"):
            key = (f[1], int(f[2]))
            vals = f[4:]
            if f[0] == "struct":
                out["struct"][key] = np.array([[int(a) for a in v.split(":")] for v in vals], np.int64).reshape(-1, 2)
            else:
                out[f[0]][key] = np.array([int(v) for v in vals], np.int64)
        elif f[0] == "approx":
            lv, up = (int(v) for v in f[1].split(":"))
            out["approx"][(lv, up, int(f[2]))] = np.array([[int(a) for a in v.split(":")] for v in f[4:]],
                                                          np.int64).reshape(-1, 2)
        elif f[0] == "precision_level":
            out["precision_level"] = (int(f[1]), int(f[2]))
        elif f[0] == "knn":
            out["knn"][int(f[1])] = [(int(v.split(":")[0]), float(v.split(":")[1])) for v in f[3:]]
        elif f[0] in ("countall", "traversal_leaves"):
            out[f[0]] = int(f[1])
        elif f[0] == "codec":
            out["codec"] = tuple(int(v) for v in f[1:])
        elif f[0] == "meta":
            out.setdefault("meta", []).append([float(f[2])] + [int(v) for v in f[3:]])
        elif f[0] in ("iterator", "kobj_mismatch"):
            raise AssertionError(line)
        elif f[0] == "kobj_ok":
            out.setdefault("kobj_ok", {})[f[1]] = int(f[2])
        elif f[0] in ("encodepoints", "getcenter", "soa", "concurrent"):
            out[f[0]] = tuple(int(v) for v in f[1:])
    return out


@pytest.mark.parametrize("enc", [A.ENC_MORTON, A.ENC_HILBERT])
def test_facade_matches_oracle(enc):
    if not os.path.exists(EXE):
        pytest.fail("facade_check is not built (make -C octrees-benchmark_amd)")
    rng = np.random.default_rng(enc)
    pts = np.vstack([rng.uniform(0, 20, size=(6000, 3)) * [1, 1, 0.3], rng.normal(size=(1000, 3)) * 0.4 + [5, 5, 2],
                     np.repeat([[12.0, 3.0, 1.0]], 20, axis=0)])
    centers = np.vstack([pts[rng.integers(0, len(pts), 25)], rng.uniform(-1, 21, size=(5, 3)) * [1, 1, 0.3]])
    radius, ml = 1.25, 16
    with tempfile.TemporaryDirectory() as d:
        pf, cf = os.path.join(d, "p.bin"), os.path.join(d, "c.bin")
        pts.astype("<f8").tofile(pf)
        centers.astype("<f8").tofile(cf)
        print(f"[gpu] facade_check enc={enc}", flush=True)
        r = subprocess.run([EXE, str(enc), str(ml), pf, cf, repr(radius)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = parse(r.stdout)

    sx, keys, perm, c, rr = O.sort_points(enc, pts)
    np.testing.assert_array_equal(got["box"], np.concatenate([c - rr, c + rr]))
    np.testing.assert_array_equal([i for i, _ in got["pt"]], perm)
    np.testing.assert_array_equal(np.array([k for _, k in got["pt"]], np.uint64), keys)
    tree = O.Tree(keys, ml, enc, c, rr)
    assert got["tree"] == (tree.nLeaf, tree.nInternal, tree.nTotal)
    names = {"sphere": O.K_SPHERE, "circle": O.K_CIRCLE, "cube": O.K_CUBE, "square": O.K_SQUARE}
    for name, kernel in names.items():
        for algo, tag in ((O.A_STRUCT, "struct"), (O.A_PRUNE, "prune"), (O.A_NEIGHBORS, "neighbors")):
            ref, sizes = O.radius(tree, sx, kernel, algo, centers, radius)
            for q in range(len(centers)):
                np.testing.assert_array_equal(got[tag][(name, q)], ref[q], err_msg=f"{tag} {name} q={q}")
    # neighborsApprox (level = the tolerance argument, as the reference's overload) and getPrecisionLevel
    for lv in range(4):
        for up in (0, 1):
            ref, _ = O.radius_approx(tree, sx, O.K_SPHERE, centers, lv, bool(up), radii=radius)
            for q in range(len(centers)):
                np.testing.assert_array_equal(got["approx"][(lv, up, q)], ref[q], err_msg=f"approx {lv}:{up} q={q}")

    def precision_level(tol):  # linear_octree.hpp:945-956, its z condition included
        rad = np.asarray(tree.radii).reshape(-1, 3)
        for i in range(tree.maxDepthSeen):
            if rad[i, 0] < tol[0] and rad[i, 1] < tol[1] and rad[i, 2] != 0 and tol[2] != 0:
                return i
        return tree.maxDepthSeen

    assert got["precision_level"] == (precision_level([radius] * 3), precision_level([radius * 50.0 / 100.0] * 3))
    bi, bd = O.knn_brute(sx, centers, 7)
    for q in range(len(centers)):
        np.testing.assert_array_equal([i for i, _ in got["knn"][q]], bi[q])
        np.testing.assert_array_equal([v for _, v in got["knn"][q]], bd[q])
    allc = O.radius_count(tree, sx, O.K_SPHERE, O.A_STRUCT, sx, radius)
    assert got["countall"] == int(allc.sum())
    assert got["traversal_leaves"] == tree.nLeaf
    k = (O.hilbert_encode if enc == A.ENC_HILBERT else O.morton_encode)(1000, 2000, 3000)
    assert got["codec"] == (k, 1000, 2000, 3000)
    # encodePoints / encodeFromPoint / getAnchorCoords give back the sorted keys; getCenter of every
    # non-empty node visited equals the tree's centre bit for bit
    assert got["encodepoints"] == (1, len(range(0, len(pts), len(pts) // 7 + 1)))
    nodes, exact = got["getcenter"]
    assert nodes > tree.nLeaf // 2 and exact == nodes
    # kernel objects (kernelFactory, kernelFactory_ptr, per-axis box radii) give the overloads' results
    assert got["kobj_ok"] == {"sphere": 8 * len(centers), "circle": 8 * len(centers), "cube": 10 * len(centers),
                              "square": 10 * len(centers)}
    # the same cloud as PointsSoA: same order, keys, tree and lists
    assert got["soa"] == (len(pts), len(pts), 1, 1, len(centers))
    # per-query calls from 8 threads, coalesced into batched launches: every caller gets its own
    # results (Struct ranges, Prune indices, kNN indices), equal to sequential calls
    checked, bad = got["concurrent"]
    assert checked > 0 and bad == 0


def test_facade_reads_las(tmp_path):
    """readPoints on a LAS file (device decode), sortPoints co-permuting the attributes, then the tree
    and a search: the oracle on the file's exactly known coordinates (power-of-two scale)."""
    from las_files import expected_attributes, write_las
    if not os.path.exists(EXE):
        pytest.fail("facade_check is not built (make -C octrees-benchmark_amd)")
    rng = np.random.default_rng(3)
    n = 4000
    X, Y, Z = rng.integers(0, 20 * 1024, n), rng.integers(0, 20 * 1024, n), rng.integers(0, 6 * 1024, n)
    path = str(tmp_path / "p.las")
    w = write_las(path, X, Y, Z, 3, minor=2, scale=(1 / 1024, 1 / 1024, 1 / 1024), offset=(100.0, 200.0, 0.0), seed=4)
    xyz = np.stack([X / 1024 + 100.0, Y / 1024 + 200.0, Z / 1024], axis=1)
    centers = xyz[rng.integers(0, n, 10)]
    cf = str(tmp_path / "c.bin")
    centers.astype("<f8").tofile(cf)
    print("[gpu] facade_check on a LAS file", flush=True)
    r = subprocess.run([EXE, str(A.ENC_HILBERT), "16", path, cf, "1.0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = parse(r.stdout)
    sx, keys, perm, c, rr = O.sort_points(A.ENC_HILBERT, xyz)
    np.testing.assert_array_equal([i for i, _ in got["pt"]], perm)
    np.testing.assert_array_equal(np.array([k for _, k in got["pt"]], np.uint64), keys)
    f, inten, psid, rgb = expected_attributes(w)
    meta = np.array(got["meta"])
    p = perm.astype(np.int64)
    np.testing.assert_array_equal(meta[:, 0], inten[p])
    np.testing.assert_array_equal(meta[:, 1:5].astype(np.int64), f[p, 0:4])
    np.testing.assert_array_equal(meta[:, 5].astype(np.int64), f[p, 4])
    np.testing.assert_array_equal(meta[:, 6].astype(np.int64) & 255, f[p, 5])
    np.testing.assert_array_equal(meta[:, 7].astype(np.int64), f[p, 6])
    np.testing.assert_array_equal(meta[:, 8].astype(np.int64), psid[p])
    np.testing.assert_array_equal(meta[:, 9:12].astype(np.int64), rgb[p])
    tree = O.Tree(keys, 16, A.ENC_HILBERT, c, rr)
    assert got["tree"] == (tree.nLeaf, tree.nInternal, tree.nTotal)
    ref, _ = O.radius(tree, sx, O.K_SPHERE, O.A_STRUCT, centers, 1.0)
    for q in range(len(centers)):
        np.testing.assert_array_equal(got["struct"][("sphere", q)], ref[q])
