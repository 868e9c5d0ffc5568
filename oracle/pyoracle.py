"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline. Nothing in the product imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ENC_MORTON, ENC_HILBERT = 1, 5
K_CIRCLE, K_SPHERE, K_SQUARE, K_CUBE = 0, 1, 2, 3
A_NEIGHBORS, A_PRUNE, A_STRUCT, A_APPROX = 1, 2, 3, 4

_P = C.c_void_p
_U64P = np.ctypeslib.ndpointer(np.uint64, flags="C")
_U32P = np.ctypeslib.ndpointer(np.uint32, flags="C")
_I64P = np.ctypeslib.ndpointer(np.int64, flags="C")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboctree_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_mbb.argtypes = [_F64P, C.c_size_t, _F64P, _F64P]
        L.orc_quantize_axis.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_quantize_axis.restype = C.c_uint64
        for f in ("orc_morton_encode", "orc_hilbert_encode"):
            getattr(L, f).argtypes = [C.c_uint64] * 3
            getattr(L, f).restype = C.c_uint64
        for f in ("orc_morton_decode", "orc_hilbert_decode"):
            getattr(L, f).argtypes = [C.c_uint64, _U64P]
        L.orc_encode.argtypes = [C.c_int, _F64P, C.c_size_t, _F64P, _F64P, _U64P]
        L.orc_sort.argtypes = [_U64P, _U32P, C.c_size_t, C.c_int]
        L.orc_gather3.argtypes = [_F64P, _U32P, C.c_size_t, _F64P]
        L.orc_build.argtypes = [_U64P, C.c_size_t, C.c_uint32, C.c_int, _F64P, _F64P]
        L.orc_build.restype = _P
        L.orc_tree_free.argtypes = [_P]
        L.orc_tree_info.argtypes = [_P, _U32P]
        for f in ("orc_tree_leaves", "orc_tree_counts", "orc_tree_layout", "orc_tree_prefixes", "orc_tree_ranges"):
            getattr(L, f).argtypes = [_P]
            getattr(L, f).restype = C.POINTER(C.c_uint64)
        L.orc_tree_offsets.argtypes = [_P]
        L.orc_tree_offsets.restype = C.POINTER(C.c_uint32)
        for f in ("orc_tree_centers", "orc_tree_radii"):
            getattr(L, f).argtypes = [_P]
            getattr(L, f).restype = C.POINTER(C.c_double)
        L.orc_radius.argtypes = [_P, _F64P, C.c_int, C.c_int, _F64P, _F64P, C.c_size_t, _P]
        L.orc_radius.restype = _P
        L.orc_radius_brute.argtypes = [_F64P, C.c_size_t, C.c_int, _F64P, _F64P, C.c_size_t]
        L.orc_radius_brute.restype = _P
        L.orc_radius_count.argtypes = [_P, _F64P, C.c_int, C.c_int, _F64P, _F64P, C.c_size_t, _P]
        L.orc_radius_count.restype = C.c_uint64
        L.orc_radius_harness.argtypes = [_P, _F64P, C.c_int, C.c_int, _F64P, _F64P, C.c_size_t]
        L.orc_radius_harness.restype = C.c_uint64
        L.orc_radius_box.argtypes = [_P, _F64P, C.c_int, C.c_int, _F64P, _F64P, C.c_size_t]
        L.orc_radius_box.restype = _P
        L.orc_radius_approx.argtypes = [_P, _F64P, C.c_int, _F64P, _F64P, C.c_void_p, C.c_size_t, C.c_uint32, C.c_int]
        L.orc_radius_approx.restype = _P
        L.orc_result_total.argtypes = [_P]
        L.orc_result_total.restype = C.c_size_t
        for f in ("orc_result_offsets", "orc_result_entries", "orc_result_sizes"):
            getattr(L, f).argtypes = [_P]
            getattr(L, f).restype = C.POINTER(C.c_uint64)
        L.orc_result_free.argtypes = [_P]
        L.orc_knn.argtypes = [_P, _F64P, C.c_size_t, _F64P, C.c_size_t, C.c_uint32, _I64P, _F64P]
        L.orc_knn.restype = C.c_uint64
        L.orc_knn_brute.argtypes = [_F64P, C.c_size_t, _F64P, C.c_size_t, C.c_uint32, _I64P, _F64P]
        L.orc_knn_brute.restype = C.c_uint64
        L.orc_max_threads.restype = C.c_int
        L.orc_ptree_build.argtypes = [_F64P, C.c_size_t, C.c_uint32, _F64P, _F64P]
        L.orc_ptree_build.restype = _P
        L.orc_ptree_free.argtypes = [_P]
        L.orc_ptree_info.argtypes = [_P, _U64P]
        L.orc_ptree_radius.argtypes = [_P, C.c_int, _F64P, C.c_size_t, C.c_double, _U64P, C.c_void_p, C.c_void_p]
        L.orc_ptree_radius.restype = C.c_uint64
        L.orc_las_header.argtypes = [C.c_char_p, _U64P, _F64P]
        L.orc_las_read.argtypes = [C.c_char_p, _F64P, _F64P, np.ctypeslib.ndpointer(np.uint8, flags="C"),
                                   np.ctypeslib.ndpointer(np.uint16, flags="C"), _U32P]
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mbb(xyz):
    xyz = _f64(xyz).reshape(-1, 3)
    c, r = np.zeros(3), np.zeros(3)
    lib().orc_mbb(xyz, xyz.shape[0], c, r)
    return c, r


def morton_encode(x, y, z):
    return lib().orc_morton_encode(int(x), int(y), int(z))


def hilbert_encode(x, y, z):
    return lib().orc_hilbert_encode(int(x), int(y), int(z))


def morton_decode(k):
    o = np.zeros(3, np.uint64)
    lib().orc_morton_decode(int(k), o)
    return tuple(int(v) for v in o)


def hilbert_decode(k):
    o = np.zeros(3, np.uint64)
    lib().orc_hilbert_decode(int(k), o)
    return tuple(int(v) for v in o)


def encode(enc, xyz, center, radii):
    xyz = _f64(xyz).reshape(-1, 3)
    keys = np.zeros(xyz.shape[0], np.uint64)
    lib().orc_encode(enc, xyz, xyz.shape[0], _f64(center), _f64(radii), keys)
    return keys


def sort(keys, nthreads=1):
    k = np.array(keys, dtype=np.uint64, copy=True)
    perm = np.zeros(k.shape[0], np.uint32)
    lib().orc_sort(k, perm, k.shape[0], nthreads)
    return k, perm


def gather3(xyz, perm):
    """xyz[perm] (n x 3), on every host core."""
    xyz = _f64(xyz).reshape(-1, 3)
    perm = np.ascontiguousarray(perm, np.uint32)
    out = np.empty((perm.shape[0], 3))
    lib().orc_gather3(xyz, perm, perm.shape[0], out)
    return out


def sort_points(enc, xyz):
    """sortPoints(points, meta) (point_encoder.hpp:484-503): returns sorted xyz, keys, perm, center, radii."""
    xyz = _f64(xyz).reshape(-1, 3)
    c, r = mbb(xyz)
    keys = encode(enc, xyz, c, r)
    sk, perm = sort(keys)
    return np.ascontiguousarray(xyz[perm.astype(np.int64)]), sk, perm, c, r


class Tree:
    def __init__(self, codes, max_leaf, enc, center, radii):
        codes = np.ascontiguousarray(codes, dtype=np.uint64)
        self.n = codes.shape[0]
        self.h = lib().orc_build(codes, self.n, max_leaf, enc, _f64(center), _f64(radii))
        if not self.h:
            raise RuntimeError("oracle build failed")
        info = np.zeros(5, np.uint32)
        lib().orc_tree_info(self.h, info)
        self.nLeaf, self.nInternal, self.nTotal, self.maxDepthSeen, self.rounds = (int(v) for v in info)

    def _arr(self, f, n, dt):
        p = getattr(lib(), f)(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

    @property
    def leaves(self):
        return self._arr("orc_tree_leaves", self.nLeaf + 1, np.uint64)

    @property
    def counts(self):
        return self._arr("orc_tree_counts", self.nLeaf, np.uint64)

    @property
    def layout(self):
        return self._arr("orc_tree_layout", self.nLeaf + 1, np.uint64)

    @property
    def prefixes(self):
        return self._arr("orc_tree_prefixes", self.nTotal, np.uint64)

    @property
    def offsets(self):
        return self._arr("orc_tree_offsets", self.nTotal + 1, np.uint32)

    @property
    def ranges(self):
        return self._arr("orc_tree_ranges", 2 * self.nTotal, np.uint64).reshape(-1, 2)

    @property
    def centers(self):
        return self._arr("orc_tree_centers", 3 * self.nTotal, np.float64).reshape(-1, 3)

    @property
    def radii(self):
        return self._arr("orc_tree_radii", 66, np.float64).reshape(22, 3)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None


def _unpack(res, nq, pairs):
    L = lib()
    total = L.orc_result_total(res)
    offs = np.ctypeslib.as_array(L.orc_result_offsets(res), shape=(nq + 1,)).astype(np.int64)
    ent = np.ctypeslib.as_array(L.orc_result_entries(res), shape=(max(total, 1),))[:total].astype(np.int64)
    sizes = np.ctypeslib.as_array(L.orc_result_sizes(res), shape=(max(nq, 1),))[:nq].astype(np.int64)
    L.orc_result_free(res)
    out = []
    for q in range(nq):
        e = ent[offs[q]:offs[q + 1]]
        out.append(e.reshape(-1, 2) if pairs else e)
    return out, sizes


def radius(tree, sorted_xyz, kernel, algo, centers, radii, stats=False):
    """Per-query outputs in the reference's emission order; Struct returns (m, 2) range arrays."""
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    radii = _f64(np.broadcast_to(np.asarray(radii, np.float64), (nq,)))
    st = np.zeros(2 * max(nq, 1), np.uint64) if stats else None
    res = lib().orc_radius(tree.h, _f64(sorted_xyz), kernel, algo, centers, radii, nq,
                           st.ctypes.data if stats else None)
    out, sizes = _unpack(res, nq, algo == A_STRUCT)
    if stats:
        return out, sizes, st.reshape(-1, 2)[:nq]
    return out, sizes


def radius_box(tree, sorted_xyz, kernel, algo, centers, radii3):
    """Cube / square with per-axis half extents; radii3 is (3,) or (nq, 3)."""
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    r3 = _f64(np.broadcast_to(np.asarray(radii3, np.float64), (nq, 3)))
    res = lib().orc_radius_box(tree.h, _f64(sorted_xyz), kernel, algo, centers, r3, nq)
    return _unpack(res, nq, algo == A_STRUCT)


def radius_approx(tree, sorted_xyz, kernel, centers, level, upper, radii=None, radii3=None):
    """neighborsApprox: Struct ranges with a depth cutoff at `level` (upper / lower bound mode)."""
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    r = _f64(np.broadcast_to(np.asarray(0.0 if radii is None else radii, np.float64), (nq,)))
    r3 = None if radii3 is None else _f64(np.broadcast_to(np.asarray(radii3, np.float64), (nq, 3)))
    res = lib().orc_radius_approx(tree.h, _f64(sorted_xyz), kernel, centers, r, None if r3 is None else r3.ctypes.data,
                                  nq, int(level), 1 if upper else 0)
    return _unpack(res, nq, True)


def radius_count(tree, sorted_xyz, kernel, algo, centers, radii):
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    radii = _f64(np.broadcast_to(np.asarray(radii, np.float64), (nq,)))
    sizes = np.zeros(max(nq, 1), np.uint64)
    lib().orc_radius_count(tree.h, _f64(sorted_xyz), kernel, algo, centers, radii, nq, sizes.ctypes.data)
    return sizes[:nq].astype(np.int64)


def radius_harness(tree, sorted_xyz, kernel, algo, centers, radii):
    """The reference harness loop (build each result, keep its size); returns the summed sizes."""
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    radii = _f64(np.broadcast_to(np.asarray(radii, np.float64), (nq,)))
    return int(lib().orc_radius_harness(tree.h, _f64(sorted_xyz), kernel, algo, centers, radii, nq))


def radius_brute(sorted_xyz, kernel, centers, radii):
    xyz = _f64(sorted_xyz).reshape(-1, 3)
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    radii = _f64(np.broadcast_to(np.asarray(radii, np.float64), (nq,)))
    res = lib().orc_radius_brute(xyz, xyz.shape[0], kernel, centers, radii, nq)
    return _unpack(res, nq, False)


def knn(tree, sorted_xyz, centers, k):
    xyz = _f64(sorted_xyz).reshape(-1, 3)
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    idx = np.zeros(nq * k, np.int64)
    d2 = np.zeros(nq * k, np.float64)
    lib().orc_knn(tree.h, xyz, xyz.shape[0], centers, nq, k, idx, d2)
    return idx.reshape(nq, k), d2.reshape(nq, k)


def knn_brute(sorted_xyz, centers, k):
    xyz = _f64(sorted_xyz).reshape(-1, 3)
    centers = _f64(centers).reshape(-1, 3)
    nq = centers.shape[0]
    idx = np.zeros(nq * k, np.int64)
    d2 = np.zeros(nq * k, np.float64)
    lib().orc_knn_brute(xyz, xyz.shape[0], centers, nq, k, idx, d2)
    return idx.reshape(nq, k), d2.reshape(nq, k)


def las_header(path):
    """orc_las_header: (version minor, format, record length, n, offset to points), scale, offset."""
    info = np.zeros(5, np.uint64)
    so = np.zeros(6)
    r = lib().orc_las_header(os.fsencode(path), info, so)
    if r:
        raise RuntimeError(f"oracle LAS header: status {r}")
    return [int(v) for v in info], so[:3].copy(), so[3:].copy()


def las_read(path):
    """orc_las_read: xyz (n x 3) and the per-point attributes as the reference's reader hands them out."""
    info, _, _ = las_header(path)
    n = info[3]
    xyz = np.zeros((max(n, 1), 3))
    inten = np.zeros(max(n, 1))
    fields = np.zeros((max(n, 1), 8), np.uint8)
    psid = np.zeros(max(n, 1), np.uint16)
    rgb = np.zeros((max(n, 1), 3), np.uint32)
    r = lib().orc_las_read(os.fsencode(path), xyz, inten, fields, psid, rgb)
    if r:
        raise RuntimeError(f"oracle LAS read: status {r}")
    return {"xyz": xyz[:n], "intensity": inten[:n], "fields": fields[:n], "psid": psid[:n], "rgb": rgb[:n]}


# ------------------------------------------------------------------ competitor baseline (reference's nanoflann)
_NF = None


def nanoflann_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libref_nanoflann.so"))


def _nf():
    global _NF
    if _NF is None:
        L = C.CDLL(os.path.join(_HERE, "_ref", "libref_nanoflann.so"))
        L.refnf_build.argtypes = [_F64P, C.c_size_t, C.c_uint32]
        L.refnf_build.restype = _P
        L.refnf_free.argtypes = [_P]
        L.refnf_radius_counts.argtypes = [_P, _F64P, C.c_size_t, C.c_double, _U64P]
        L.refnf_knn.argtypes = [_P, _F64P, C.c_size_t, C.c_uint32, _I64P, _F64P]
        _NF = L
    return _NF


class NanoflannTree:
    """The reference's vendored nanoflann KD-tree (competitor baseline), leaf size maxPointsLeaf."""

    def __init__(self, xyz, leaf):
        self.xyz = _f64(xyz).reshape(-1, 3)  # kept alive: the tree reads the caller's points
        self.h = _nf().refnf_build(self.xyz, self.xyz.shape[0], leaf)

    def radius_counts(self, centers, radius):
        centers = _f64(centers).reshape(-1, 3)
        out = np.zeros(centers.shape[0], np.uint64)
        _nf().refnf_radius_counts(self.h, centers, centers.shape[0], float(radius), out)
        return out

    def knn(self, centers, k):
        centers = _f64(centers).reshape(-1, 3)
        idx = np.zeros((centers.shape[0], k), np.int64)
        d2 = np.zeros((centers.shape[0], k))
        _nf().refnf_knn(self.h, centers, centers.shape[0], k, idx, d2)
        return idx, d2

    def __del__(self):
        if getattr(self, "h", None):
            _nf().refnf_free(self.h)
            self.h = None


# ------------------------------------------------------------------ competitor baseline (reference's pointer octree, restated)
class PointerOctree:
    """The reference's pointer-based Octree (inc/structures/octree.hpp, src/octree.cpp) restated in the
    oracle (orc_ptree_*): insertion-built with leaves of at most max_leaf points, searched by its
    neighbors() (boxOverlap descent, isInside per point); a CPU competitor baseline."""

    def __init__(self, xyz, max_leaf, center, radii):
        self.xyz = _f64(xyz).reshape(-1, 3)  # kept alive: the tree reads the caller's points
        self.h = lib().orc_ptree_build(self.xyz, self.xyz.shape[0], max_leaf, _f64(center), _f64(radii))
        info = np.zeros(3, np.uint64)
        lib().orc_ptree_info(self.h, info)
        self.n_leaf, self.n_internal, self.depth = (int(v) for v in info)

    def radius_counts(self, kernel, centers, radius):
        centers = _f64(centers).reshape(-1, 3)
        sizes = np.zeros(centers.shape[0], np.uint64)
        lib().orc_ptree_radius(self.h, kernel, centers, centers.shape[0], float(radius), sizes, None, None)
        return sizes

    def radius(self, kernel, centers, radius):
        """Index lists in the reference's emission order."""
        centers = _f64(centers).reshape(-1, 3)
        sizes = self.radius_counts(kernel, centers, radius)
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        out = np.zeros(max(int(offs[-1]), 1), np.uint64)
        lib().orc_ptree_radius(self.h, kernel, centers, centers.shape[0], float(radius), sizes.copy(),
                               offs.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        return [out[offs[q]:offs[q + 1]].astype(np.int64) for q in range(centers.shape[0])]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ptree_free(self.h)
            self.h = None
