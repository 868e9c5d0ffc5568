"""CPU: the facade's host-side parts against the oracle, without a GPU (facade_host_check).

- Kernel objects (kernelFactory / kernelFactory_ptr, inc/kernels/kernel_factory.hpp:14-41): their
  isInside, boxOverlap and boxIntersect drive the reference's neighborsPrune / neighbors traversals
  (linear_octree.hpp:746-816) over the oracle's tree; the lists must equal the oracle's for all four
  kernels, box kernels with per-axis radii included, and isInside over every point must equal the
  oracle's brute force.
- PointsAoS / PointsSoA (point_containers.hpp:32-163) hold the same points, and NeighborSet iterates
  both alike.
- MortonEncoder2D keys equal the reference's libmorton (tests/golden/morton_ref.npz) and a bit-by-bit
  interleave; HilbertEncoder2D decodes its own keys, walks the curve in unit steps from (0, 0) and has
  the level-1 digit order (0,0) (0,1) (1,1) (1,0) of hilbert_encoder_2d.hpp:29-34; NoEncoding and the
  encoder factory report the reference's constants (point_encoder_factory.hpp:12-44).
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "octrees-benchmark_amd", "facade_host_check")
NAMES = {"sphere": O.K_SPHERE, "circle": O.K_CIRCLE, "cube": O.K_CUBE, "square": O.K_SQUARE}


def interleave2(a, b):
    k = 0
    for j in range(32):
        k |= ((a >> j) & 1) << (2 * j) | ((b >> j) & 1) << (2 * j + 1)
    return k


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("facade_host_check is not built (make -C octrees-benchmark_amd)")
    d = str(tmp_path_factory.mktemp("facade_host"))
    rng = np.random.default_rng(17)
    pts = np.vstack([rng.uniform(0, 20, (4000, 3)) * [1, 1, 0.3], rng.normal(size=(600, 3)) * 0.4 + [5, 5, 2],
                     np.repeat([[12.0, 3.0, 1.0]], 30, axis=0),
                     np.round(rng.uniform(0, 20, (400, 3)) * [1, 1, 0.3] * 4) / 4])  # points on a 0.25 m grid
    sx, keys, perm, c, r = O.sort_points(O.ENC_HILBERT, pts)
    tree = O.Tree(keys, 16, O.ENC_HILBERT, c, r)
    q = np.vstack([sx[rng.integers(0, len(sx), 14)], rng.uniform(-1, 21, (3, 3)) * [1, 1, 0.3],
                   [[12.0, 3.0, 1.0], [5.0, 5.0, 2.0]]])
    sx.astype("<f8").tofile(os.path.join(d, "pts.f64"))
    q.astype("<f8").tofile(os.path.join(d, "q.f64"))
    np.asarray(tree.offsets[:tree.nTotal], "<u4").tofile(os.path.join(d, "off.u32"))
    np.asarray(tree.ranges, "<u8").tofile(os.path.join(d, "rng.u64"))
    np.asarray(tree.centers, "<f8").tofile(os.path.join(d, "ctr.f64"))
    np.asarray(tree.radii, "<f8").tofile(os.path.join(d, "rad.f64"))
    g = np.load(os.path.join(ROOT, "tests", "golden", "morton_ref.npz"))
    np.asarray(g["xy2"], "<u8").tofile(os.path.join(d, "pairs.u64"))
    radius = 1.25
    p = subprocess.run([EXE, d, repr(radius)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    out = {"lists": {}, "pair": [], "enc2d": [], "encoder": {}}
    for line in p.stdout.splitlines():
        f = line.split()
        if f[0] in ("prune", "neighbors", "brute", "boxprune"):
            out["lists"][(f[0], f[1], int(f[2]))] = np.array([int(v) for v in f[4:]], np.int64)
        elif f[0] == "pair":
            out["pair"].append((int(f[1]), int(f[2])))
        elif f[0] == "enc2d":
            out["enc2d"].append(tuple(int(v) for v in f[1:]))
        elif f[0] == "encoder":
            out["encoder"][int(f[1])] = f[2:]
        elif f[0] in ("containers", "neighborset", "hilbert2d_walk", "axis"):
            out[f[0]] = [int(v) for v in f[1:]]
        else:
            raise AssertionError(f"unexpected line: {line}")
    return out, sx, tree, q, radius, g


@pytest.mark.parametrize("name", list(NAMES))
def test_kernel_objects_match_oracle(run, name):
    out, sx, tree, q, radius, _ = run
    k = NAMES[name]
    prune, _ = O.radius(tree, sx, k, O.A_PRUNE, q, radius)
    nbrs, _ = O.radius(tree, sx, k, O.A_NEIGHBORS, q, radius)
    brute, _ = O.radius_brute(sx, k, q, radius)
    for i in range(len(q)):
        np.testing.assert_array_equal(out["lists"][("prune", name, i)], prune[i], err_msg=f"prune {name} q={i}")
        np.testing.assert_array_equal(out["lists"][("neighbors", name, i)], nbrs[i], err_msg=f"neighbors {name} q={i}")
        np.testing.assert_array_equal(out["lists"][("brute", name, i)], np.sort(brute[i]), err_msg=f"brute {name} q={i}")


@pytest.mark.parametrize("name", ["cube", "square"])
def test_box_kernel_objects_with_axis_radii(run, name):
    out, sx, tree, q, radius, _ = run
    ref, _ = O.radius_box(tree, sx, NAMES[name], O.A_PRUNE, q, [radius, 0.6 * radius, 1.4 * radius])
    for i in range(len(q)):
        np.testing.assert_array_equal(out["lists"][("boxprune", name, i)], ref[i], err_msg=f"{name} q={i}")


def test_containers_and_neighbor_set(run):
    out = run[0]
    n = len(run[1])
    assert out["containers"] == [n, n, n]
    size, visited, agree, ended = out["neighborset"]
    assert size == visited == agree and size > 0 and ended == 1


def test_morton2d_matches_reference_libmorton(run):
    out, *_, g = run
    keys = np.array([m for m, _ in out["pair"]], np.uint64)
    np.testing.assert_array_equal(keys, g["keys2"])
    for (a, b), k in zip(g["xy2"][:200], keys[:200]):
        assert interleave2(int(a), int(b)) == int(k)


def test_hilbert2d_curve(run):
    out = run[0]
    # level-1 digits and the round trips (decode(encode) = id) of the edge pairs
    first = {(a, b): h for a, b, m, h, mok, hok in out["enc2d"]}
    assert first[(0, 0)] == 0 and first[(0, 1)] == 1 and first[(1, 1)] == 2 and first[(1, 0)] == 3
    assert all(mok == 1 and hok == 1 for *_, mok, hok in out["enc2d"])
    # 256 consecutive keys: 255 unit steps, all inside the 16 x 16 corner, each re-encoding to its key
    assert out["hilbert2d_walk"] == [255, 256]
    # a bijection on the full-depth keys of the random pairs
    hk = [h for _, h in out["pair"]]
    assert len(set(hk)) == len(set(map(tuple, run[5]["xy2"].tolist())))
    # axis X keeps (y, z), axis Y keeps (x, z)
    kx, x0, y0, z0, ky, x1, y1, z1 = out["axis"]
    assert kx == interleave2(5, 9) and (x0, y0, z0) == (0, 5, 9) and (x1, y1, z1) == (5, 0, 9)


def test_encoder_factory_constants(run):
    e = run[0]["encoder"]
    assert e[0][:1] == ["0"] and "NoEncoding" in " ".join(e[0])
    # decode(encode(3, 4, 5)) keeps the two coordinates of the encoder's plane
    plane = {0: ["0", "4", "5"], 1: ["3", "0", "5"], 2: ["3", "4", "0"]}
    for t in (2, 3, 4, 6, 7, 8):
        f = e[t]
        assert int(f[0]) == t
        assert f[-3:] == plane[(t - 2) % 4], (t, f)
    m = " ".join(e[2])
    assert "MortonEncoder2D Axis X mort_2d_axis_X 31 2 4611686018427387904 0" in m
    assert "HilbertEncoder2D Axis Z hilb_2d_axis_Z 31 2 9223372036854775808 0" in " ".join(e[8])
    assert " ".join(e[1]).startswith("1 MortonEncoder3D mort 21 1 9223372036854775808 1")
    assert " ".join(e[5]).startswith("5 HilbertEncoder3D hilb 21 1 9223372036854775808 1")
    assert e[9][0] == "0"  # unknown values give NoEncoding
    # MortonEncoder2D's eps keeps the reference's expression 1.0f / (1 << 31), negative under C++20
    assert float(e[2][-4]) == -(2.0 ** -31) and float(e[8][-4]) == 2.0 ** -31
