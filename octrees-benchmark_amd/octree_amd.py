"""Python binding of the gfx950 octree library (liboctree_amd.so) through its C ABI
(include/octree_amd.h), plus the deterministic synthetic-cloud generators (liboctree_synth.so).

The library must be built in place first (`make -C octrees-benchmark_amd`, or
`__graft_entry__.build()`); loading fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboctree_amd.so")
SYNTH_PATH = os.path.join(HERE, "liboctree_synth.so")

ENC_MORTON, ENC_HILBERT = 1, 5
K_CIRCLE, K_SPHERE, K_SQUARE, K_CUBE = 0, 1, 2, 3
A_NEIGHBORS, A_PRUNE, A_STRUCT = 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
STAGES = ("bbox", "encode", "sort", "gather", "build", "search", "knn")

# every symbol declared in include/octree_amd.h
EXPORTS = (
    "oct_create", "oct_destroy", "oct_last_error", "oct_synchronize", "oct_version", "oct_set_points",
    "oct_set_points_dev", "oct_bbox", "oct_encode", "oct_sort", "oct_sort_points", "oct_set_sorted", "oct_build",
    "oct_get_tree", "oct_radius", "oct_radius_approx", "oct_radius_fetch", "oct_radius_dev", "oct_knn", "oct_knn_dev", "oct_knn_fetch", "oct_autotune_knn", "oct_tuned_knn", "oct_stage_times",
    "oct_set_tuning", "oct_autotune_radius", "oct_tuned_radius", "oct_radius_box", "oct_encode_coords",
    "oct_decode_keys", "oct_batch_create", "oct_batch_destroy", "oct_batch_last_error", "oct_batch_add",
    "oct_batch_build", "oct_batch_info", "oct_batch_stage_times", "oct_batch_radius", "oct_batch_knn",
    "oct_batch_build_device", "oct_batch_set_points", "oct_batch_cloud_sorted", "oct_batch_cloud_tree",
    "oct_batch_queries", "oct_batch_pipeline", "oct_batch_pipeline_results", "oct_batch_autotune_knn",
    "oct_las_read_header", "oct_las_load", "oct_las_fetch", "oct_gather_records", "oct_set_points_soa",
    "oct_set_sorted_soa", "oct_knn_checksum", "oct_copy_bandwidth", "oct_radius_lists_noalloc",
)

# oct_point_meta (per-point LAS attributes, PointMetadata)
POINT_META = np.dtype([("intensity", "<f8"), ("return_number", "u1"), ("number_of_returns", "u1"),
                       ("scan_direction", "u1"), ("edge_of_flight_line", "u1"), ("classification", "u1"),
                       ("scan_angle_rank", "i1"), ("user_data", "u1"), ("pad0", "u1"), ("point_source_id", "<u2"),
                       ("pad1", "<u2"), ("r", "<u4"), ("g", "<u4"), ("b", "<u4")])
assert POINT_META.itemsize == 32


class LasHeader(C.Structure):
    _fields_ = [("version_major", C.c_uint8), ("version_minor", C.c_uint8), ("point_format", C.c_uint8),
                ("pad0", C.c_uint8), ("record_length", C.c_uint16), ("header_size", C.c_uint16),
                ("offset_to_points", C.c_uint32), ("n_vlr", C.c_uint32), ("n_points", C.c_uint64),
                ("scale", C.c_double * 3), ("offset", C.c_double * 3), ("min", C.c_double * 3), ("max", C.c_double * 3)]

    def as_dict(self):
        d = {}
        for f, _ in self._fields_:
            v = getattr(self, f)
            d[f] = list(v) if isinstance(v, C.Array) else v
        return d


class OctError(RuntimeError):
    pass


class TreeInfo(C.Structure):
    _fields_ = [("n_leaf", C.c_uint32), ("n_internal", C.c_uint32), ("n_total", C.c_uint32),
                ("max_depth_seen", C.c_uint32), ("max_points_leaf", C.c_uint32), ("n_levels", C.c_uint32),
                ("build_ms", C.c_double)]


class SearchStats(C.Structure):
    _fields_ = [("total_neighbors", C.c_uint64), ("points_tested", C.c_uint64), ("nodes_classified", C.c_uint64),
                ("rechecked", C.c_uint64), ("kernel_ms", C.c_double), ("class_slots", C.c_uint64),
                ("test_slots", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_LIB = None
_SYNTH = None
_HIP = None
_vp = C.c_void_p


class DeviceArray:
    """A device copy of a numpy array for the device-pointer entry points, allocated with the HIP
    runtime the library loaded; ptr(offset) is the address of element `offset`; close() frees it."""

    def __init__(self, arr):
        global _HIP
        lib()
        if _HIP is None:
            _HIP = C.CDLL("libamdhip64.so.7")
            _HIP.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
            _HIP.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
            _HIP.hipFree.argtypes = [_vp]
        self.arr = np.ascontiguousarray(arr)
        self.p = _vp()
        if _HIP.hipMalloc(C.byref(self.p), max(self.arr.nbytes, 16)) != 0:
            raise OctError("hipMalloc failed")
        if self.arr.nbytes and _HIP.hipMemcpy(self.p, self.arr.ctypes.data, self.arr.nbytes, 1) != 0:
            self.close()
            raise OctError("hipMemcpy to the device failed")

    def ptr(self, offset=0):
        return _vp(self.p.value + int(offset) * self.arr.itemsize)

    def close(self):
        if self.p:
            _HIP.hipFree(self.p)
            self.p = _vp()


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OctError(f"{LIB_PATH} is not built (run `make -C octrees-benchmark_amd`); no CPU fallback exists")
        L = C.CDLL(LIB_PATH)
        L.oct_create.argtypes = [C.POINTER(_vp), C.c_int]
        L.oct_destroy.argtypes = [_vp]
        L.oct_last_error.argtypes = [_vp]
        L.oct_last_error.restype = C.c_char_p
        L.oct_version.restype = C.c_char_p
        L.oct_synchronize.argtypes = [_vp]
        L.oct_set_points.argtypes = [_vp, _vp, C.c_size_t, C.c_size_t]
        L.oct_set_points_dev.argtypes = [_vp, _vp, C.c_size_t, C.c_size_t]
        L.oct_bbox.argtypes = [_vp, _vp, _vp]
        L.oct_encode.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
        L.oct_sort.argtypes = [_vp, _vp, _vp]
        L.oct_sort_points.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp]
        L.oct_set_sorted.argtypes = [_vp, _vp, C.c_size_t, C.c_size_t, _vp, C.c_int, _vp, _vp]
        L.oct_build.argtypes = [_vp, C.c_uint32, C.POINTER(TreeInfo)]
        L.oct_get_tree.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.oct_radius.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_double, C.c_size_t, C.c_int, _vp,
                                 C.POINTER(SearchStats)]
        L.oct_radius_fetch.argtypes = [_vp, _vp, _vp, C.c_size_t]
        L.oct_radius_dev.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_double, C.c_size_t, _vp,
                                     C.POINTER(SearchStats)]
        L.oct_knn.argtypes = [_vp, C.c_int, _vp, _vp, C.c_size_t, C.c_uint32, _vp, _vp, C.POINTER(SearchStats)]
        L.oct_knn_dev.argtypes = [_vp, C.c_int, _vp, C.c_size_t, C.c_size_t, C.c_uint32, _vp, _vp, C.POINTER(SearchStats)]
        L.oct_knn_fetch.argtypes = [_vp, C.c_size_t, C.c_size_t, _vp, _vp]
        L.oct_knn_checksum.argtypes = [_vp, _vp, _vp, C.c_size_t, C.c_uint32, _vp]
        L.oct_copy_bandwidth.argtypes = [_vp, C.c_size_t, C.POINTER(C.c_double)]
        L.oct_radius_lists_noalloc.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_double, C.c_size_t, _vp, _vp,
                                               _vp, C.c_size_t, C.POINTER(C.c_uint64)]
        L.oct_set_points_soa.argtypes = [_vp, _vp, _vp, _vp, C.c_size_t]
        L.oct_set_sorted_soa.argtypes = [_vp, _vp, _vp, _vp, C.c_size_t, _vp, C.c_int, _vp, _vp]
        L.oct_radius_approx.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_double, _vp, C.c_size_t, C.c_uint32,
                                        C.c_int, C.c_int, _vp, C.POINTER(SearchStats)]
        L.oct_batch_create.argtypes = [C.POINTER(_vp), C.c_int]
        L.oct_batch_destroy.argtypes = [_vp]
        L.oct_batch_last_error.argtypes = [_vp]
        L.oct_batch_last_error.restype = C.c_char_p
        L.oct_batch_add.argtypes = [_vp, _vp, C.c_size_t, C.c_size_t, C.c_int, C.c_uint32, C.POINTER(C.c_int)]
        L.oct_batch_build.argtypes = [_vp]
        L.oct_batch_build_device.argtypes = [_vp, C.c_int, C.POINTER(C.c_double)]
        L.oct_batch_set_points.argtypes = [_vp, C.c_int, _vp, C.c_size_t]
        L.oct_batch_cloud_sorted.argtypes = [_vp, C.c_int, _vp, _vp]
        L.oct_batch_cloud_tree.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp]
        L.oct_gather_records.argtypes = [_vp, _vp, _vp, C.c_size_t, C.c_int]
        L.oct_las_read_header.argtypes = [C.c_char_p, C.POINTER(LasHeader)]
        L.oct_las_load.argtypes = [_vp, C.c_char_p, C.POINTER(LasHeader)]
        L.oct_las_fetch.argtypes = [_vp, _vp, _vp]
        L.oct_batch_queries.argtypes = [_vp, _vp, _vp, _vp, C.c_size_t, _vp, _vp, _vp, C.c_size_t, C.c_uint32]
        L.oct_batch_pipeline.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.oct_batch_pipeline_results.argtypes = [_vp, _vp, _vp, _vp]
        L.oct_batch_autotune_knn.argtypes = [_vp, C.c_int, _vp, _vp]
        L.oct_batch_stage_times.argtypes = [_vp, _vp]
        L.oct_batch_info.argtypes = [_vp, C.c_int, C.POINTER(C.c_size_t), C.POINTER(TreeInfo)]
        L.oct_batch_radius.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_size_t, _vp, C.POINTER(SearchStats)]
        L.oct_batch_knn.argtypes = [_vp, C.c_int, _vp, _vp, _vp, C.c_size_t, C.c_uint32, _vp, _vp, C.POINTER(SearchStats)]
        L.oct_autotune_knn.argtypes = [_vp, C.c_int, C.c_uint32, C.c_size_t, C.POINTER(C.c_int), _vp]
        L.oct_tuned_knn.argtypes = [_vp, C.c_int, C.c_uint32, C.POINTER(C.c_int)]
        L.oct_stage_times.argtypes = [_vp, _vp]
        L.oct_set_tuning.argtypes = [_vp, C.c_int, C.c_int]
        L.oct_autotune_radius.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_size_t,
                                          C.POINTER(C.c_int), _vp]
        L.oct_tuned_radius.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.oct_radius_box.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_size_t, C.c_int, _vp,
                                     C.POINTER(SearchStats)]
        L.oct_encode_coords.argtypes = [_vp, C.c_int, _vp, C.c_size_t, _vp]
        L.oct_decode_keys.argtypes = [_vp, C.c_int, _vp, C.c_size_t, _vp]
        _LIB = L
    return _LIB


def synth():
    global _SYNTH
    if _SYNTH is None:
        if not os.path.exists(SYNTH_PATH):
            raise OctError(f"{SYNTH_PATH} is not built (run `make -C octrees-benchmark_amd`)")
        S = C.CDLL(SYNTH_PATH)
        S.oct_synth_uniform.argtypes = [C.c_uint64, C.c_size_t] + [C.c_double] * 6 + [_vp]
        S.oct_synth_lidar_tile.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_int,
                                           C.c_int, C.c_double, C.c_double, _vp]
        S.oct_synth_indices.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, _vp]
        S.oct_synth_mls_sweep.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, C.c_double, _vp]
        _SYNTH = S
    return _SYNTH


def _ptr(a):
    return None if a is None else a.ctypes.data


def las_header(path):
    """oct_las_read_header: the public header block of a LAS file (no context, no GPU)."""
    h = LasHeader()
    rc = lib().oct_las_read_header(os.fsencode(path), C.byref(h))
    if rc != 0:
        raise OctError(f"oct_las_read_header({path}) failed with status {rc}")
    return h.as_dict()


# ------------------------------------------------------------------ workloads
def uniform_box(n, seed=1, size=(100.0, 100.0, 20.0), offset=(0.0, 0.0, 0.0)):
    """Uniform cloud in a box on a LAS millimetre grid (the 1 M-point plain workload uses 100x100x20 m)."""
    xyz = np.empty((n, 3), np.float64)
    synth().oct_synth_uniform(seed, n, *size, *offset, _ptr(xyz))
    return xyz


def lidar_tile(n, seed=2, extent=1000.0, offset=(651000.0, 6861000.0), n_buildings=None, n_blobs=None,
               frac_building=0.15, frac_veg=0.20):
    """Synthetic airborne-LiDAR tile: sinusoid terrain + box buildings + Gaussian vegetation blobs.
    Defaults scale building/blob counts with the area (2000 buildings and 4000 blobs per km^2)."""
    area = extent * extent / 1e6
    nb = int(round(2000 * area)) if n_buildings is None else n_buildings
    nv = int(round(4000 * area)) if n_blobs is None else n_blobs
    xyz = np.empty((n, 3), np.float64)
    synth().oct_synth_lidar_tile(seed, n, extent, offset[0], offset[1], max(nb, 1), max(nv, 1), frac_building,
                                 frac_veg, _ptr(xyz))
    return xyz


def mls_sweep(n, seed=5, length=5000.0, offset=(651000.0, 6861000.0)):
    """Synthetic mobile-laser-scan sweep along a ~length-metre smooth trajectory: ground returns with
    range density ~ 1/d^2 on [1, 80] m, facade planes, pole cylinders, 0.1 % outliers."""
    xyz = np.empty((n, 3), np.float64)
    synth().oct_synth_mls_sweep(seed, n, length, offset[0], offset[1], _ptr(xyz))
    return xyz


def random_indices(nq, n, seed=42):
    out = np.empty(nq, np.uint32)
    synth().oct_synth_indices(seed, nq, n, _ptr(out))
    return out


class BatchWorkload:
    """The batched workload: `clouds` independent clouds with sizes log-uniform in [1e4, 4e6], rescaled
    to ~`total` points; odd clouds uniform boxes (5 pts/m^3 shape, z extent a fifth), even clouds
    airborne tiles at 50 pts/m^2; each at a random origin in [-1e5, 1e5]. `per` centres per cloud
    drawn uniformly from its points (sorted-order indices), a radius per query uniform in [0.5, 5] m
    rounded to float, a fraction `knn_frac` of the queries kNN with k uniform in {8..256}.
    Deterministic in `seed`; clouds are generated on demand by cloud(i)."""

    def __init__(self, seed=4, clouds=256, total=200e6, per=50_000, knn_frac=0.2):
        rng = np.random.default_rng(seed)
        C_ = clouds
        sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), C_))
        self.sizes = np.maximum(1e4, sizes * (total / sizes.sum())).astype(np.int64)
        self.offsets = [rng.uniform(-1e5, 1e5, 3) for _ in range(C_)]
        self.per = per
        self.cloud_of = np.repeat(np.arange(C_, dtype=np.uint32), per)
        self.idx = (rng.uniform(0.0, 1.0, C_ * per) * np.repeat(self.sizes, per)).astype(np.uint32)
        self.is_knn = rng.uniform(0.0, 1.0, C_ * per) < knn_frac
        self.radii = rng.uniform(0.5, 5.0, C_ * per).astype(np.float32).astype(np.float64)
        self.k = rng.integers(8, 257, C_ * per).astype(np.uint32)
        self.rq = np.nonzero(~self.is_knn)[0]
        self.kq = np.nonzero(self.is_knn)[0]

    @property
    def count(self):
        return len(self.sizes)

    @property
    def points(self):
        return int(self.sizes.sum())

    def cloud(self, i):
        n, off = int(self.sizes[i]), self.offsets[i]
        if i % 2:
            L = float(n) ** (1.0 / 3.0)
            return uniform_box(n, seed=1000 + i, size=(L, L, L / 5.0), offset=tuple(off))
        return lidar_tile(n, seed=1000 + i, extent=1000.0 * (n / 50e6) ** 0.5, offset=(off[0], off[1]))

    def radius_queries(self):
        """(cloud, index, radius) of the radius queries"""
        return self.cloud_of[self.rq], self.idx[self.rq], self.radii[self.rq]

    def knn_queries(self):
        """(cloud, index, k) of the kNN queries"""
        return self.cloud_of[self.kq], self.idx[self.kq], self.k[self.kq]


# ------------------------------------------------------------------ context
class Batch:
    """Many independent clouds: each is sorted and built as on its own, then all are searched together
    (one launch per search kind). Query centres are (cloud, index in that cloud's sorted order)."""

    def __init__(self, device=0):
        self.L = lib()
        h = _vp()
        rc = self.L.oct_batch_create(C.byref(h), device)
        if rc != 0:
            raise OctError(f"oct_batch_create(device={device}) failed with status {rc}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.oct_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.oct_batch_last_error(self.h)
            raise OctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def add(self, xyz, encoder=ENC_HILBERT, max_points_leaf=128):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        cid = C.c_int(-1)
        self._chk(self.L.oct_batch_add(self.h, _ptr(xyz), xyz.shape[0], 24, encoder, max_points_leaf, C.byref(cid)),
                  "oct_batch_add")
        return cid.value

    def build(self):
        self._chk(self.L.oct_batch_build(self.h), "oct_batch_build")

    def build_device(self, graph=True):
        """Every cloud built at once on the device (one HIP graph when `graph`); returns device ms."""
        ms = C.c_double(0.0)
        self._chk(self.L.oct_batch_build_device(self.h, 1 if graph else 0, C.byref(ms)), "oct_batch_build_device")
        return ms.value

    def set_points(self, cloud, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self._chk(self.L.oct_batch_set_points(self.h, cloud, _ptr(xyz), 24), "oct_batch_set_points")

    def cloud_sorted(self, cloud):
        n, _ = self.info(cloud)
        keys = np.empty(n, np.uint64)
        perm = np.empty(n, np.uint32)
        self._chk(self.L.oct_batch_cloud_sorted(self.h, cloud, _ptr(keys), _ptr(perm)), "oct_batch_cloud_sorted")
        return keys, perm

    def cloud_tree(self, cloud):
        nt = self.info(cloud)[1]["n_total"]
        pre = np.empty(nt, np.uint64)
        off = np.empty(nt, np.uint32)
        rng = np.empty((nt, 2), np.uint64)
        cen = np.empty((nt, 3), np.float64)
        rad = np.empty((22, 3), np.float64)
        self._chk(self.L.oct_batch_cloud_tree(self.h, cloud, _ptr(pre), _ptr(off), _ptr(rng), _ptr(cen), _ptr(rad)),
                  "oct_batch_cloud_tree")
        return {"prefixes": pre, "offsets": off, "ranges": rng, "centers": cen, "radii": rad}

    def stage_times(self):
        ms = np.zeros(8)
        self._chk(self.L.oct_batch_stage_times(self.h, _ptr(ms)), "oct_batch_stage_times")
        return dict(zip(STAGES, ms[:len(STAGES)].tolist()))

    def info(self, cloud):
        n = C.c_size_t(0)
        ti = TreeInfo()
        self._chk(self.L.oct_batch_info(self.h, cloud, C.byref(n), C.byref(ti)), "oct_batch_info")
        return n.value, {f: getattr(ti, f) for f, _ in ti._fields_}

    def queries(self, r_cloud=(), r_idx=(), r_radius=(), k_cloud=(), k_idx=(), k_k=(), kmax=None):
        """Device-resident query set of the search pass (oct_batch_queries)."""
        rc_ = np.ascontiguousarray(r_cloud, np.uint32)
        ri = np.ascontiguousarray(r_idx, np.uint32)
        rr = np.ascontiguousarray(r_radius, np.float64)
        kc = np.ascontiguousarray(k_cloud, np.uint32)
        ki = np.ascontiguousarray(k_idx, np.uint32)
        kk = np.ascontiguousarray(k_k, np.uint32)
        kmax = (int(kk.max()) if len(kk) else 0) if kmax is None else kmax
        self._nq = (len(rc_), len(kc), kmax)
        self._chk(self.L.oct_batch_queries(self.h, _ptr(rc_), _ptr(ri), _ptr(rr), len(rc_), _ptr(kc), _ptr(ki), _ptr(kk),
                                           len(kc), kmax), "oct_batch_queries")

    def pipeline(self, kernel=0, algo=3, precision=1, graph=True):
        """Device build + grouped searches (one HIP graph when `graph`); returns device ms."""
        ms = C.c_double(0.0)
        self._chk(self.L.oct_batch_pipeline(self.h, kernel, algo, precision, 1 if graph else 0, C.byref(ms)),
                  "oct_batch_pipeline")
        return ms.value

    def pipeline_results(self, knn_d2=True):
        nr, nk, kmax = self._nq
        counts = np.empty(nr, np.uint32)
        idx = np.empty((nk, kmax), np.int32)
        d2 = np.empty((nk, kmax), np.float64) if knn_d2 else None
        self._chk(self.L.oct_batch_pipeline_results(self.h, _ptr(counts), _ptr(idx), _ptr(d2) if knn_d2 else None),
                  "oct_batch_pipeline_results")
        return counts, idx, d2

    def autotune_knn(self, precision=1):
        lanes = np.zeros(7, np.int32)
        ms = np.zeros(21)
        self._chk(self.L.oct_batch_autotune_knn(self.h, precision, _ptr(lanes), _ptr(ms)), "oct_batch_autotune_knn")
        return lanes.tolist(), ms.reshape(7, 3)

    def radius(self, kernel, algo, precision, cloud, idx, radii):
        cloud = np.ascontiguousarray(cloud, np.uint32)
        idx = np.ascontiguousarray(idx, np.uint32)
        radii = np.ascontiguousarray(radii, np.float64)
        nq = cloud.shape[0]
        counts = np.empty(nq, np.uint32)
        st = SearchStats()
        self._chk(self.L.oct_batch_radius(self.h, kernel, algo, precision, _ptr(cloud), _ptr(idx), _ptr(radii), nq,
                                          _ptr(counts), C.byref(st)), "oct_batch_radius")
        return counts, st.as_dict()

    def knn(self, precision, cloud, idx, k, kmax=None):
        cloud = np.ascontiguousarray(cloud, np.uint32)
        idx = np.ascontiguousarray(idx, np.uint32)
        k = np.ascontiguousarray(k, np.uint32)
        nq = cloud.shape[0]
        kmax = int(k.max()) if kmax is None else kmax
        oi = np.empty((nq, kmax), np.int32)
        od = np.empty((nq, kmax), np.float64)
        st = SearchStats()
        self._chk(self.L.oct_batch_knn(self.h, precision, _ptr(cloud), _ptr(idx), _ptr(k), nq, kmax, _ptr(oi), _ptr(od),
                                       C.byref(st)), "oct_batch_knn")
        return oi, od, st.as_dict()


class Octree:
    """One device context: points, keys, sorted cloud, tree and search buffers live on the GPU."""

    def __init__(self, device=0):
        self.L = lib()
        h = _vp()
        rc = self.L.oct_create(C.byref(h), device)
        if rc != 0:
            raise OctError(f"oct_create(device={device}) failed with status {rc}")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.oct_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.oct_last_error(self.h)
            raise OctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- H1-H3
    def gather_records(self, records):
        """records (n rows of any fixed-size dtype, input order) reordered by the last sort's permutation."""
        records = np.ascontiguousarray(records)
        out = np.empty_like(records)
        rb = records.dtype.itemsize * (int(np.prod(records.shape[1:])) if records.ndim > 1 else 1)
        self._chk(self.L.oct_gather_records(self.h, _ptr(records), _ptr(out), rb, 0), "oct_gather_records")
        return out

    def gather_records_dev(self, in_ptr, out_ptr, record_bytes):
        """Device-buffer form (e.g. torch tensors' data_ptr())."""
        self._chk(self.L.oct_gather_records(self.h, in_ptr, out_ptr, record_bytes, 1), "oct_gather_records")

    def las_load(self, path):
        """The LAS file's points become the context's cloud (decoded on the device); returns its header."""
        h = LasHeader()
        self._chk(self.L.oct_las_load(self.h, os.fsencode(path), C.byref(h)), "oct_las_load")
        self.current = None
        self.n = self._las_n = int(h.n_points)
        return h.as_dict()

    def las_fetch(self, meta=True):
        n = self._las_n
        xyz = np.empty((n, 3), np.float64)
        m = np.empty(n, POINT_META) if meta else None
        self._chk(self.L.oct_las_fetch(self.h, _ptr(xyz), _ptr(m) if meta else None), "oct_las_fetch")
        return xyz, m

    def set_points(self, xyz):
        """xyz: (n, 3) packed points, or (n, 4) rows laid out like the reference's 32-byte Point."""
        xyz = np.ascontiguousarray(xyz, np.float64)
        cols = 4 if (xyz.ndim == 2 and xyz.shape[1] == 4) else 3
        xyz = xyz.reshape(-1, cols)
        self.n = xyz.shape[0]
        self._chk(self.L.oct_set_points(self.h, _ptr(xyz), self.n, 8 * cols), "oct_set_points")

    def bbox(self):
        c, r = np.zeros(3), np.zeros(3)
        self._chk(self.L.oct_bbox(self.h, _ptr(c), _ptr(r)), "oct_bbox")
        return c, r

    def encode(self, encoder, center, radii, fetch=True):
        keys = np.empty(self.n, np.uint64) if fetch else None
        c, r = np.ascontiguousarray(center, np.float64), np.ascontiguousarray(radii, np.float64)
        self._chk(self.L.oct_encode(self.h, encoder, _ptr(c), _ptr(r), _ptr(keys)), "oct_encode")
        return keys

    def sort(self, fetch=True):
        perm = np.empty(self.n, np.uint32) if fetch else None
        keys = np.empty(self.n, np.uint64) if fetch else None
        self._chk(self.L.oct_sort(self.h, _ptr(perm), _ptr(keys)), "oct_sort")
        return perm, keys

    def sort_points(self, encoder, fetch=True):
        c, r = np.zeros(3), np.zeros(3)
        perm = np.empty(self.n, np.uint32) if fetch else None
        keys = np.empty(self.n, np.uint64) if fetch else None
        self._chk(self.L.oct_sort_points(self.h, encoder, _ptr(c), _ptr(r), _ptr(perm), _ptr(keys)), "oct_sort_points")
        return c, r, perm, keys

    def set_sorted(self, xyz, codes, encoder, center, radii):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        codes = np.ascontiguousarray(codes, np.uint64)
        self.n = xyz.shape[0]
        c, r = np.ascontiguousarray(center, np.float64), np.ascontiguousarray(radii, np.float64)
        self._chk(self.L.oct_set_sorted(self.h, _ptr(xyz), self.n, 24, _ptr(codes), encoder, _ptr(c), _ptr(r)),
                  "oct_set_sorted")

    # ---- H4
    def build(self, max_points_leaf=128):
        info = TreeInfo()
        self._chk(self.L.oct_build(self.h, max_points_leaf, C.byref(info)), "oct_build")
        self.info = {f: getattr(info, f) for f, _ in info._fields_}
        return self.info

    def get_tree(self):
        nt = self.info["n_total"]
        pre = np.empty(nt, np.uint64)
        off = np.empty(nt, np.uint32)
        rng = np.empty((nt, 2), np.uint64)
        cen = np.empty((nt, 3), np.float64)
        rad = np.empty((22, 3), np.float64)
        self._chk(self.L.oct_get_tree(self.h, _ptr(pre), _ptr(off), _ptr(rng), _ptr(cen), _ptr(rad)), "oct_get_tree")
        return {"prefixes": pre, "offsets": off, "ranges": rng, "centers": cen, "radii": rad}

    # ---- H5
    def radius(self, kernel, algo, precision, centers=None, idx=None, radii=None, radius=0.0, nq=None, lists=False):
        if centers is not None:
            centers = np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
            nq = centers.shape[0]
        elif idx is not None:
            idx = np.ascontiguousarray(idx, np.uint32)
            nq = idx.shape[0]
        elif nq is None:
            nq = self.n
        if radii is not None:
            radii = np.ascontiguousarray(np.broadcast_to(np.asarray(radii, np.float64), (nq,)))
        counts = np.empty(nq, np.uint32)
        st = SearchStats()
        self._chk(self.L.oct_radius(self.h, kernel, algo, precision, _ptr(centers), _ptr(idx), _ptr(radii),
                                    float(radius), nq, 1 if lists else 0, _ptr(counts), C.byref(st)), "oct_radius")
        if not lists:
            return counts, st.as_dict()
        return self._fetch_lists(algo, nq, counts, st)

    def radius_box(self, kernel, algo, radii3, centers=None, idx=None, nq=None, lists=False):
        """Cube / square searches with per-axis half extents: radii3 is one (3,) triple or (nq, 3)."""
        if centers is not None:
            centers = np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
            nq = centers.shape[0]
        elif idx is not None:
            idx = np.ascontiguousarray(idx, np.uint32)
            nq = idx.shape[0]
        elif nq is None:
            nq = self.n
        r3 = np.ascontiguousarray(radii3, np.float64)
        per_query = r3.ndim == 2
        counts = np.empty(nq, np.uint32)
        st = SearchStats()
        self._chk(self.L.oct_radius_box(self.h, kernel, algo, _ptr(centers), _ptr(idx), _ptr(r3) if per_query else None,
                                        None if per_query else _ptr(r3), nq, 1 if lists else 0, _ptr(counts),
                                        C.byref(st)), "oct_radius_box")
        if not lists:
            return counts, st.as_dict()
        return self._fetch_lists(algo, nq, counts, st)

    def encode_coords(self, encoder, xyz):
        xyz = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
        keys = np.empty(xyz.shape[0], np.uint64)
        self._chk(self.L.oct_encode_coords(self.h, encoder, _ptr(xyz), xyz.shape[0], _ptr(keys)), "oct_encode_coords")
        return keys

    def decode_keys(self, encoder, keys):
        keys = np.ascontiguousarray(keys, np.uint64)
        xyz = np.empty((keys.shape[0], 3), np.uint32)
        self._chk(self.L.oct_decode_keys(self.h, encoder, _ptr(keys), keys.shape[0], _ptr(xyz)), "oct_decode_keys")
        return xyz

    def _fetch_lists(self, algo, nq, counts, st):
        offs = np.empty(nq + 1, np.uint64)
        self._chk(self.L.oct_radius_fetch(self.h, _ptr(offs), None, 0), "oct_radius_fetch(offsets)")
        words = int(offs[-1]) * (2 if algo == A_STRUCT else 1)
        ent = np.empty(max(words, 1), np.uint32)
        self._chk(self.L.oct_radius_fetch(self.h, _ptr(offs), _ptr(ent), ent.shape[0]), "oct_radius_fetch")
        out = []
        for q in range(nq):
            a, b = int(offs[q]), int(offs[q + 1])
            if algo == A_STRUCT:
                out.append(ent[2 * a:2 * b].reshape(-1, 2).astype(np.int64))
            else:
                out.append(ent[a:b].astype(np.int64))
        return counts, out, st.as_dict()

    def radius_approx(self, kernel, precision, level, upper, centers=None, idx=None, radius=0.0, radius3=None, nq=None,
                      lists=False):
        """neighborsApprox: Struct with a depth cutoff at `level`; upper-bound or lower-bound mode."""
        if centers is not None:
            centers = np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
            nq = centers.shape[0]
        elif idx is not None:
            idx = np.ascontiguousarray(idx, np.uint32)
            nq = idx.shape[0]
        elif nq is None:
            nq = self.n
        r3 = None if radius3 is None else np.ascontiguousarray(radius3, np.float64).reshape(3)
        counts = np.empty(nq, np.uint32)
        st = SearchStats()
        self._chk(self.L.oct_radius_approx(self.h, kernel, precision, _ptr(centers), _ptr(idx), None, float(radius), _ptr(r3),
                                           nq, int(level), 1 if upper else 0, 1 if lists else 0, _ptr(counts), C.byref(st)),
                  "oct_radius_approx")
        if not lists:
            return counts, st.as_dict()
        return self._fetch_lists(A_STRUCT, nq, counts, st)

    def radius_csr(self, kernel, algo, precision, idx, radius):
        """Result lists of a search over sorted-cloud indices as CSR: counts, offsets (nq + 1) and the flat
        entries (point indices, or [first, last) pairs for neighborsStruct), without per-query arrays."""
        idx = np.ascontiguousarray(idx, np.uint32)
        nq = idx.shape[0]
        counts = np.empty(nq, np.uint32)
        st = SearchStats()
        self._chk(self.L.oct_radius(self.h, kernel, algo, precision, None, _ptr(idx), None, float(radius), nq, 1,
                                    _ptr(counts), C.byref(st)), "oct_radius")
        offs = np.empty(nq + 1, np.uint64)
        self._chk(self.L.oct_radius_fetch(self.h, _ptr(offs), None, 0), "oct_radius_fetch(offsets)")
        words = int(offs[-1]) * (2 if algo == A_STRUCT else 1)
        ent = np.empty(max(words, 1), np.uint32)
        self._chk(self.L.oct_radius_fetch(self.h, _ptr(offs), _ptr(ent), ent.shape[0]), "oct_radius_fetch")
        return counts, offs, ent[:words], st.as_dict()

    def radius_dev(self, kernel, algo, precision, radius, nq=None, stats=False, idx_ptr=None):
        """All-points search on the device-resident sorted cloud (or, with idx_ptr, the nq sorted-cloud
        indices in that device buffer as centres); counts stay on the device."""
        nq = self.n if nq is None else nq
        st = SearchStats() if stats else None
        self._chk(self.L.oct_radius_dev(self.h, kernel, algo, precision, idx_ptr, float(radius), nq, None,
                                        C.byref(st) if stats else None), "oct_radius_dev")
        return st.as_dict() if stats else None

    # ---- H6
    def knn(self, k, precision=PREC_FP64, centers=None, idx=None, nq=None):
        if centers is not None:
            centers = np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
            nq = centers.shape[0]
        elif idx is not None:
            idx = np.ascontiguousarray(idx, np.uint32)
            nq = idx.shape[0]
        elif nq is None:
            nq = self.n
        oi = np.empty((nq, k), np.int32)
        od = np.empty((nq, k), np.float64)
        st = SearchStats()
        self._chk(self.L.oct_knn(self.h, precision, _ptr(centers), _ptr(idx), nq, k, _ptr(oi), _ptr(od), C.byref(st)),
                  "oct_knn")
        return oi, od, st.as_dict()

    def knn_dev(self, k, precision=PREC_FP32, first=0, nq=None, stats=False, idx_ptr=None, d2_ptr=None):
        """All-points kNN for sorted points first .. first + nq - 1; results (int32 indices, FP32 d^2)
        stay on the device until knn_fetch, or go to the caller's device buffers idx_ptr / d2_ptr
        (addresses of nq x k int32 / float32 arrays on the context's device, both given)."""
        nq = self.n - first if nq is None else nq
        st = SearchStats() if stats else None
        own = idx_ptr is None or d2_ptr is None
        self._chk(self.L.oct_knn_dev(self.h, precision, None, first, nq, k, None if own else idx_ptr,
                                     None if own else d2_ptr, C.byref(st) if stats else None), "oct_knn_dev")
        if own:
            self._knn_shape = (nq, k)
        return st.as_dict() if stats else None

    def knn_fetch(self, q0=0, nq=None):
        total, k = self._knn_shape
        nq = total - q0 if nq is None else nq
        oi = np.empty((nq, k), np.int32)
        od = np.empty((nq, k), np.float32)
        self._chk(self.L.oct_knn_fetch(self.h, q0, nq, _ptr(oi), _ptr(od)), "oct_knn_fetch")
        return oi, od

    def knn_checksum(self, nq=None, k=None, idx_ptr=None, d2_ptr=None):
        """(checksum, bad rows, index sum) of kNN results on the device: the kept results of the last
        knn_dev (first nq rows), or the caller's device buffers idx_ptr / d2_ptr of nq x k entries."""
        if idx_ptr is None:
            total, kk = self._knn_shape
            nq, k = (total if nq is None else nq), kk
        out = np.zeros(3, np.uint64)
        self._chk(self.L.oct_knn_checksum(self.h, idx_ptr, d2_ptr, nq, k, _ptr(out)), "oct_knn_checksum")
        return int(out[0]), int(out[1]), int(out[2])

    def copy_bandwidth(self, nbytes=1 << 30):
        """Read + write GB/s of the library's streaming copy over nbytes (best of 5 timed runs)."""
        ms = C.c_double(0.0)
        self._chk(self.L.oct_copy_bandwidth(self.h, nbytes, C.byref(ms)), "oct_copy_bandwidth")
        return 2 * nbytes / (ms.value * 1e-3) / 1e9, ms.value

    def set_points_soa(self, x, y, z):
        """The cloud as three coordinate arrays (PointsSoA)."""
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        self._chk(self.L.oct_set_points_soa(self.h, _ptr(x), _ptr(y), _ptr(z), x.shape[0]), "oct_set_points_soa")
        self.n = x.shape[0]

    def set_remap(self, on):
        """Count-mode leaf remapping of sparsely wanted leaves (None: default, on)."""
        self._chk(self.L.oct_set_tuning(self.h, 4, -1 if on is None else (1 if on else 0)), "oct_set_tuning")

    def set_sort(self, mode):
        """How oct_sort sorts: 1 partition + 2048-key LDS sorts (default, also None), 2 the same with
        4096-key LDS sorts, 0 the device-wide LSD pass per 8-bit digit."""
        self._chk(self.L.oct_set_tuning(self.h, 5, -1 if mode is None else int(mode)), "oct_set_tuning")

    def set_knn_lanes(self, lanes):
        """Lanes per query of the kNN kernel (None: autotuned / default)."""
        self._chk(self.L.oct_set_tuning(self.h, 3, -1 if lanes is None else int(lanes)), "oct_set_tuning")

    def autotune_knn(self, k, precision=PREC_FP32, sample_nq=0):
        best = C.c_int(0)
        ms = np.zeros(3)
        self._chk(self.L.oct_autotune_knn(self.h, precision, k, sample_nq, C.byref(best), _ptr(ms)), "oct_autotune_knn")
        return best.value, ms

    def tuned_knn(self, k, precision=PREC_FP32):
        lanes = C.c_int(0)
        self._chk(self.L.oct_tuned_knn(self.h, precision, k, C.byref(lanes)), "oct_tuned_knn")
        return lanes.value

    def set_group(self, g):
        """Queries per wave of the count-mode radius kernel (0, 16, 32, 64; -1 = tuned/default)."""
        self._chk(self.L.oct_set_tuning(self.h, 1, g), "oct_set_tuning")

    def set_mfma(self, on):
        """The OCT_TUNE_RADIUS_MFMA knob: the point tests run on the vector ALUs, so None and False are
        accepted and True raises OctError (no matrix-core form is built)."""
        self._chk(self.L.oct_set_tuning(self.h, 2, -1 if on is None else (1 if on else 0)), "oct_set_tuning")

    def tuned_radius(self, kernel, algo, precision):
        """Queries per wave a count-mode search would use now."""
        g, m = C.c_int(0), C.c_int(0)
        self._chk(self.L.oct_tuned_radius(self.h, kernel, algo, precision, C.byref(g), C.byref(m)), "oct_tuned_radius")
        return g.value

    def autotune_radius(self, kernel, algo, precision, radius, sample_nq=1 << 20):
        best = C.c_int(0)
        ms = np.zeros(4)
        self._chk(self.L.oct_autotune_radius(self.h, kernel, algo, precision, float(radius), sample_nq, C.byref(best),
                                             _ptr(ms)), "oct_autotune_radius")
        return best.value, dict(zip((0, 16, 32, 64), ms.tolist()))

    def synchronize(self):
        self._chk(self.L.oct_synchronize(self.h), "oct_synchronize")

    def stage_times(self):
        ms = np.zeros(8)
        self._chk(self.L.oct_stage_times(self.h, _ptr(ms)), "oct_stage_times")
        return dict(zip(STAGES, ms[:len(STAGES)].tolist()))
