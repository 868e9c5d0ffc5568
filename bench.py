#!/usr/bin/env python3
"""Benchmark: all-points radius search on the 50 M-point synthetic airborne-LiDAR tile, one MI355X.

A step is one pass of the hot path's search over the device-resident tree: 50 M sphere queries
(r = 2.5 m, neighborsStruct semantics, per-query neighbour counts), FP32 + FP64 recheck. The same
run also times the plain FP64 search, encode + sort (+ gather) and the build, checks that both
search precisions give identical counts for every query and identical neighbour lists to the CPU
oracle on a sample, and times the CPU oracle (the reference algorithm restated) on a bounded
query sample on the host's cores.

Prints ONE JSON line.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "octrees-benchmark_amd"))

import octree_amd as A  # noqa: E402

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3


# sources that the all-points count kernel is compiled from: a committed counter profile of that
# kernel applies to this build only when their digest matches
SEARCH_SOURCES = ("search.hip", "search_common.hpp", "device.hpp", "tree_common.hpp", "internal.hpp")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def source_digest(names=SEARCH_SOURCES):
    h = hashlib.sha256()
    d = os.path.join(ROOT, "octrees-benchmark_amd", "csrc")
    for f in names:
        h.update(f.encode())
        h.update(open(os.path.join(d, f), "rb").read())
    return h.hexdigest()[:16]


def cpu_model():
    """The host CPU's model name, as lscpu reports it (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_of(fn, reps):
    """One warm-up call, then the median wall time of `reps` timed calls (and the last result)."""
    out = fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), out


def sync_all(ctx):
    ctx.synchronize()


def profile(args):
    """Small fixed workloads for rocprofv3 runs (kernel statistics and counters)."""
    if args.profile == "copy":
        ctx = A.Octree(0)
        for _ in range(3):
            gbs, ms = ctx.copy_bandwidth(1 << 30)
            log(f"[profile] copy 1 GiB: {ms:.3f} ms = {gbs:.0f} GB/s (read + write)")
        ctx.close()
        return
    n = args.n
    extent = args.extent if n == 50_000_000 else args.extent * (n / 50_000_000) ** 0.5
    ctx = A.Octree(0)
    ctx.set_points(A.lidar_tile(n, seed=2, extent=extent))
    if args.profile == "sort":
        for _ in range(3):
            ctx.sort_points(A.ENC_HILBERT)
            ctx.build(args.max_leaf)
            log(f"[profile] {ctx.stage_times()}")
        # the streaming copy in the same process: the practical ceiling the stages are read against
        for _ in range(2):
            gbs, ms = ctx.copy_bandwidth(1 << 30)
            log(f"[profile] copy 1 GiB: {ms:.3f} ms = {gbs:.0f} GB/s (read + write)")
        ctx.close()
        return
    kind, _, spec = args.profile.partition(":")
    if kind == "knn" and spec:  # 'knn:10' -> one all-points kNN per precision, default lanes per query
        ctx.sort_points(A.ENC_HILBERT, fetch=False)
        ctx.build(args.max_leaf)
        for prec in (A.PREC_FP32, A.PREC_FP64):
            ctx.knn_dev(int(spec), prec)
            ctx.synchronize()
            log(f"[profile] knn k={spec} precision {prec}: {ctx.stage_times()['knn']:.2f} ms")
        ctx.close()
        return
    if kind != "search" or not spec:
        sys.exit(f"unknown --profile {args.profile!r}")
    ctx.sort_points(A.ENC_HILBERT)
    ctx.build(args.max_leaf)
    for v in spec.split(","):
        prec = A.PREC_FP64 if v[0] == "d" else A.PREC_FP32
        ctx.set_group(int(v[1:]))
        ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, args.radius)
        ctx.synchronize()
        log(f"[profile] {v}: {ctx.stage_times()['search']:.2f} ms")
    ctx.close()


def mls_stress(args):
    """Stress workload: a mobile-laser-scan sweep (default 400 M points along a 5 km trajectory),
    Morton against Hilbert, all-points r = 0.5 m sphere counts in both precisions, and neighborsStruct
    range lists (CSR) for a 10 M-centre subset. One JSON line; not the benchmark line."""
    n = args.n if args.n != 50_000_000 else 400_000_000
    radius = 0.5
    t0 = time.time()
    xyz = A.mls_sweep(n)
    log(f"[mls] generated {n} points in {time.time() - t0:.1f} s")
    ctx = A.Octree(0)
    ctx.set_points(xyz)
    out = {"workload": f"synthetic mobile-laser-scan sweep, {n} FP64 points along a 5 km trajectory, maxLeaf "
                       f"{args.max_leaf}, all-points sphere r={radius} m counts, neighborsStruct CSR for 10 M centres",
           "points": n, "radius": radius, "encoders": {}}
    per_point = {}
    for enc, name in ((A.ENC_MORTON, "morton"), (A.ENC_HILBERT, "hilbert")):
        first = None  # the first sort of these points (it may try the partition sort before the LSD passes)
        for rep in range(2):
            c, r, perm, keys = ctx.sort_points(enc, fetch=(rep == 1))
            if rep == 0:
                first = ctx.stage_times()["sort"]
        st = ctx.stage_times()
        info = ctx.build(args.max_leaf)
        bst = ctx.stage_times()["build"]
        e = {"encode_ms": st["encode"], "sort_ms": st["sort"], "first_sort_ms": first, "gather_ms": st["gather"],
             "bbox_ms": st["bbox"],
             "build_ms": bst, "tree": {k: info[k] for k in ("n_leaf", "n_internal", "n_total", "max_depth_seen")},
             "keys_sorted": bool(np.all(keys[1:] >= keys[:-1]))}
        log(f"[mls] {name}: sort {st['sort']:.2f} ms build {bst:.2f} ms tree {info}")
        cnt = {}
        for pname, prec in (("fp32", A.PREC_FP32), ("fp64", A.PREC_FP64)):
            ctx.autotune_radius(A.K_SPHERE, A.A_STRUCT, prec, radius, sample_nq=1 << 22)
            ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, radius)
            ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, radius)
            ms = ctx.stage_times()["search"]
            cnt[pname], stt = ctx.radius(A.K_SPHERE, A.A_STRUCT, prec, radius=radius, nq=n)
            e[f"search_{pname}_ms"] = ms
            e[f"{pname}_queries_per_s"] = n / (ms * 1e-3)
            e["total_neighbors"] = stt["total_neighbors"]
            log(f"[mls] {name} {pname}: {ms:.1f} ms, {stt['total_neighbors']} neighbours")
        e["fp32_eq_fp64_all_queries"] = bool(np.array_equal(cnt["fp32"], cnt["fp64"]))
        e["fp32_vs_fp64_speedup"] = e["search_fp64_ms"] / e["search_fp32_ms"]
        orig = np.empty(n, np.uint32)
        orig[perm.astype(np.int64)] = cnt["fp32"]
        per_point[name] = orig
        del cnt, keys
        # neighborsStruct lists for 10 M random centres: CSR of [first, last) ranges
        nq = min(10_000_000, n)
        idx = A.random_indices(nq, n, seed=5)
        counts, offs, ent, stl = ctx.radius_csr(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, idx, radius)
        rng = ent.reshape(-1, 2).astype(np.int64)
        cs = np.concatenate([[0], np.cumsum(rng[:, 1] - rng[:, 0])])
        lens = cs[offs[1:].astype(np.int64)] - cs[offs[:-1].astype(np.int64)]
        e["csr"] = {"centres": nq, "ranges": int(offs[-1]), "neighbors": int(counts.astype(np.int64).sum()),
                    "kernel_ms": stl["kernel_ms"], "range_lengths_eq_counts": bool(np.array_equal(lens, counts))}
        log(f"[mls] {name} csr: {e['csr']}")
        del rng, ent, offs, counts, perm
        out["encoders"][name] = e
    out["counts_eq_across_encoders"] = bool(np.array_equal(per_point["morton"], per_point["hilbert"]))
    print(json.dumps(out), flush=True)
    ctx.close()


def batch_workload(args):
    """Batched workload: 256 independent clouds (sizes log-uniform in [1e4, 4e6], rescaled to ~200 M
    points; half uniform boxes at 5 pts/m^3, half terrain + building tiles at 50 pts/m^2; each at a
    random origin in [-1e5, 1e5]), 50 k random centres per cloud, radius per query uniform in
    [0.5, 5] m (float), 20 % of the queries kNN with k uniform in {8..256}. The whole sequence
    (per-cloud boxes, keys, segmented sort, forest build, query grouping, work-queue count and kNN
    kernels) runs as one HIP graph; the same clouds built one by one on per-cloud contexts and
    searched by the per-call batch API check it. One JSON line of its own."""
    rng = np.random.default_rng(4)
    C = 256
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), C))
    sizes = np.maximum(1e4, sizes * (200e6 / sizes.sum())).astype(np.int64)
    t0 = time.time()
    b = A.Batch(0)   # built and searched on the device in one HIP graph
    bp = A.Batch(0)  # the same clouds built one by one (per-cloud contexts), for comparison
    clouds = []
    for i, n in enumerate(sizes):
        n = int(n)
        off = rng.uniform(-1e5, 1e5, 3)
        if i % 2:
            L = float(n) ** (1.0 / 3.0)
            xyz = A.uniform_box(n, seed=1000 + i, size=(L, L, L / 5.0), offset=tuple(off))
        else:
            xyz = A.lidar_tile(n, seed=1000 + i, extent=1000.0 * (n / 50e6) ** 0.5, offset=(off[0], off[1]))
        b.add(xyz, A.ENC_HILBERT, args.max_leaf)
        bp.add(xyz, A.ENC_HILBERT, args.max_leaf)
        clouds.append(xyz if i < 3 else None)
    gen_s = time.time() - t0
    npts = int(sizes.sum())
    log(f"[batch] {C} clouds, {npts} points generated and uploaded in {gen_s:.1f} s")
    per = 50_000
    cloud = np.repeat(np.arange(C, dtype=np.uint32), per)
    idx = (rng.uniform(0.0, 1.0, C * per) * np.repeat(sizes, per)).astype(np.uint32)
    is_knn = rng.uniform(0.0, 1.0, C * per) < 0.2
    radii = rng.uniform(0.5, 5.0, C * per).astype(np.float32).astype(np.float64)
    kk = rng.integers(8, 257, C * per).astype(np.uint32)
    rq, kq = np.nonzero(~is_knn)[0], np.nonzero(is_knn)[0]
    steps = max(1, args.steps)

    # build only (bbox, keys, segmented sort, forest), one graph
    first_ms = b.build_device(graph=True)
    bms = [b.build_device(graph=True) for _ in range(steps)]
    build_ms = float(np.median(bms))
    log(f"[batch] device build graph: first {first_ms:.1f} ms, replays {build_ms:.2f} ms device")
    # the whole pass, one graph: build + query grouping + work-queue count and kNN kernels
    b.queries(cloud[rq], idx[rq], radii[rq], cloud[kq], idx[kq], kk[kq], kmax=256)
    lanes, tune_ms = b.autotune_knn(A.PREC_FP32)
    log(f"[batch] autotuned kNN lanes per query by k class (16, 32, ... 1024): {lanes}")
    out = {"workload": f"{C} clouds, {npts} points (log-uniform 1e4..4e6), {per} centres each, radius "
                       f"U[0.5, 5] m, 20 % kNN with k U{{8..256}}, Hilbert, maxLeaf {args.max_leaf}",
           "clouds": C, "points": npts, "radius_queries": int(len(rq)), "knn_queries": int(len(kq)),
           "build_graph": {"replay_device_ms": build_ms, "points_per_s": npts / (build_ms * 1e-3)},
           "knn_lanes_by_class": lanes}
    res = {}
    for name, prec in (("fp32", A.PREC_FP32), ("fp64", A.PREC_FP64)):
        b.pipeline(A.K_SPHERE, A.A_STRUCT, prec, graph=True)
        pm, pw = [], []
        for _ in range(steps):
            t = time.perf_counter()
            pm.append(b.pipeline(A.K_SPHERE, A.A_STRUCT, prec, graph=True))
            pw.append(time.perf_counter() - t)
        rep = float(np.median(pm))
        res[name] = b.pipeline_results()
        out[f"pipeline_graph_{name}"] = {"replay_device_ms": rep, "replay_wall_ms": float(np.median(pw)) * 1e3,
                                         "search_ms": rep - build_ms,
                                         "queries_per_s": (len(rq) + len(kq)) / (rep * 1e-3),
                                         "neighbors": int(res[name][0].astype(np.int64).sum())}
        log(f"[batch] {name}: whole pass in one graph {rep:.1f} ms device ({float(np.median(pw)) * 1e3:.1f} ms wall): "
            f"build {build_ms:.1f} + grouping and searches {rep - build_ms:.1f}")
    # the per-call batch API on the same forest: one launch per kind (kNN: per k class), host grouping
    for name, prec in (("fp32", A.PREC_FP32),):
        t = time.perf_counter()
        counts, _ = b.radius(A.K_SPHERE, A.A_STRUCT, prec, cloud[rq], idx[rq], radii[rq])
        wall = time.perf_counter() - t
        dev = b.stage_times()["search"]
        t = time.perf_counter()
        ki, kd, st = b.knn(prec, cloud[kq], idx[kq], kk[kq], 256)
        kwall = time.perf_counter() - t
        out[f"per_call_{name}"] = {"radius_device_ms": dev, "radius_wall_s": wall, "knn_device_ms": st["kernel_ms"],
                                   "knn_wall_s": kwall}
        log(f"[batch] per-call {name}: radius {dev:.1f} ms device ({wall:.2f} s wall), knn {st['kernel_ms']:.1f} ms "
            f"device ({kwall:.2f} s wall)")
        same_per_call = bool(np.array_equal(counts, res["fp32"][0])) and bool(np.array_equal(ki, res["fp32"][1]))
    # the same clouds built one by one
    t0 = time.time()
    bp.build()
    out["build_per_cloud_s"] = time.time() - t0
    cp, _ = bp.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, cloud[rq], idx[rq], radii[rq])
    ksub = kq[:100_000]
    pki, pkd, _ = bp.knn(A.PREC_FP32, cloud[ksub], idx[ksub], kk[ksub], 256)
    same_per_cloud = bool(np.array_equal(cp, res["fp32"][0])) and bool(np.array_equal(pki, res["fp32"][1][:len(ksub)]))
    bp.close()
    log(f"[batch] per-cloud build {out['build_per_cloud_s']:.2f} s")
    # exact against the oracle on the first three clouds (each on its own)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    ok = True
    for c in range(3):
        xyz = clouds[c]
        sx, keys, perm, cc, rr = O.sort_points(A.ENC_HILBERT, xyz)
        tree = O.Tree(keys, args.max_leaf, A.ENC_HILBERT, cc, rr)
        rsel = np.nonzero(cloud[rq] == c)[0][:300]
        for q in rsel:
            qq = rq[q]
            ref = O.radius_count(tree, sx, A.K_SPHERE, A.A_STRUCT, sx[[int(idx[qq])]], float(radii[qq]))[0]
            ok &= int(res["fp32"][0][q]) == int(ref)
        ksel = np.nonzero(cloud[kq] == c)[0][:100]
        bi, bd = O.knn_brute(sx, sx[idx[kq[ksel]].astype(np.int64)], 256)
        for row, q in enumerate(ksel):
            m = int(kk[kq[q]])
            ok &= bool(np.array_equal(res["fp32"][1][q, :m].astype(np.int64), bi[row, :m]))
    out["verify"] = {"fp32_eq_fp64": bool(np.array_equal(res["fp32"][0], res["fp64"][0]))
                     and bool(np.array_equal(res["fp32"][1], res["fp64"][1])),
                     "pipeline_eq_per_call": same_per_call, "pipeline_eq_per_cloud_build": same_per_cloud,
                     "oracle_sample_three_clouds": bool(ok)}
    print(json.dumps(out), flush=True)
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=50_000_000)
    ap.add_argument("--extent", type=float, default=1000.0)
    ap.add_argument("--radius", type=float, default=2.5)
    ap.add_argument("--max-leaf", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--knn-k", default="10,50,250,1000",
                    help="k values of the all-points kNN measured beside the search, comma-separated (0: skip); "
                         "k <= 50 runs every point, k = 250 the first 10 M and k = 1000 the first 2 M SFC positions")
    ap.add_argument("--cpu-reps", type=int, default=2, help="timed repeats of each CPU-baseline stage (median)")
    ap.add_argument("--workload", default="tile50m", choices=("tile50m", "mls400m", "batch256"),
                    help="tile50m: the benchmark line (default); mls400m: the mobile-laser-scan stress run; batch256: "
                         "256 independent clouds searched together (each prints one JSON line of its own)")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-competitor", action="store_true",
                    help="skip the reference's competitor, its vendored nanoflann KD-tree, on the CPU")
    ap.add_argument("--las", default=None, help="run the benchmark line on the points of this (uncompressed) LAS file "
                                                 "instead of the synthetic tile")
    ap.add_argument("--profile", default=None,
                    help="profiling driver instead of the benchmark: 'search:v64,d64' (one all-points search per "
                         "configuration: v = FP32 + recheck, d = FP64, then queries per "
                         "wave), 'knn:10' (one all-points kNN per precision), 'sort' (bbox + encode + sort + gather "
                         "+ build, 3 times, then the 1 GiB copy kernel) or 'copy' (the 1 GiB copy kernel alone, the "
                         "counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(the multi-GPU sharding is a design note, not an executable path)", file=sys.stderr)
        sys.exit(2)
    if args.profile:
        profile(args)
        return
    if args.workload == "mls400m":
        mls_stress(args)
        return
    if args.workload == "batch256":
        batch_workload(args)
        return

    ctx = A.Octree(0)
    las = None
    if args.las:
        # a LAS file instead of the synthetic tile: header parsed on the host, records decoded on the device
        t0 = time.perf_counter()
        hdr = ctx.las_load(args.las)
        ingest_s = time.perf_counter() - t0
        xyz, _ = ctx.las_fetch(meta=False)
        n = xyz.shape[0]
        extent = float(max(hdr["max"][0] - hdr["min"][0], hdr["max"][1] - hdr["min"][1]))
        las = {"file": os.path.basename(args.las), "points": n, "format": hdr["point_format"],
               "version": f"{hdr['version_major']}.{hdr['version_minor']}", "ingest_wall_s": ingest_s,
               "ingest_mb_per_s": n * hdr["record_length"] / ingest_s / 1e6}
        log(f"[bench] LAS {args.las}: {n} points (format {hdr['point_format']}) read and decoded in {ingest_s:.2f} s")
    else:
        n = args.n
        extent = args.extent if args.n == 50_000_000 else args.extent * (n / 50_000_000) ** 0.5
        t0 = time.time()
        xyz = A.lidar_tile(n, seed=2, extent=extent)
        log(f"[bench] generated {n} points over {extent:.0f} m in {time.time() - t0:.1f} s")
        ctx.set_points(xyz)
    # encode + sort + build (an untimed warm-up that also fetches the keys and permutation, then 5
    # repetitions timed with device events; each stage's median is reported)
    sort_ms, enc_ms, bbox_ms, gather_ms, build_ms = [], [], [], [], []
    for rep in range(6):
        c, r, perm, keys = ctx.sort_points(A.ENC_HILBERT, fetch=(rep == 0))
        if rep == 0:
            keys0, perm0 = keys, perm
        info = ctx.build(args.max_leaf)
        st = ctx.stage_times()
        bbox_ms.append(st["bbox"]), enc_ms.append(st["encode"]), sort_ms.append(st["sort"])
        gather_ms.append(st["gather"]), build_ms.append(st["build"])
    med = lambda v: float(np.median(v[1:]))  # noqa: E731
    log(f"[bench] tree: {info}")
    log(f"[bench] bbox {med(bbox_ms):.3f} ms encode {med(enc_ms):.3f} ms sort {med(sort_ms):.3f} ms "
        f"gather {med(gather_ms):.3f} ms build {med(build_ms):.3f} ms")

    K, W = args.steps, args.warmup
    results = {}
    tuned = {}
    # Each precision gets its own best tile shape, tuned on the full workload: FP32 + recheck and
    # plain FP64.
    for name, prec in (("fp32", A.PREC_FP32), ("fp64", A.PREC_FP64)):
        best, cand = ctx.autotune_radius(A.K_SPHERE, A.A_STRUCT, prec, args.radius, sample_nq=n)
        g = ctx.tuned_radius(A.K_SPHERE, A.A_STRUCT, prec)
        tuned[name] = {"queries_per_wave": g, "candidate_ms": cand}
        log(f"[bench] autotune {name}: G={g} candidates(ms)={cand}")
        for _ in range(W):
            ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, args.radius)
        sync_all(ctx)
        t = time.perf_counter()
        kms = []
        for _ in range(K):
            ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, args.radius)
            kms.append(ctx.stage_times()["search"])
        sync_all(ctx)
        wall = (time.perf_counter() - t) / K
        stats = ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, prec, args.radius, stats=True)
        results[name] = {"wall_s": wall, "kernel_ms": float(np.mean(kms)), "stats": stats}
        log(f"[bench] {name} step {wall * 1e3:.2f} ms kernel {np.mean(kms):.2f} ms stats {stats}")

    fp32, fp64 = results["fp32"], results["fp64"]

    # kNN (KNNV2) on the same tile: every point as a query, in chunks of consecutive SFC positions where
    # the int32 indices + FP32 squared distances of all queries would not fit the device (k = 1000 over
    # 50 M points would be 400 GB); each precision with its autotuned lanes-per-query shape. The kernel's
    # counters and the oracle sample come from the first chunk.
    knn = {}
    kvals = [int(v) for v in str(args.knn_k).split(",") if int(v) > 0]
    for k in kvals:
        chunk = min(n, n if k <= 50 else (10_000_000 if k <= 250 else 2_000_000))
        ent = {"k": k, "queries": n, "chunk": chunk}
        for name, prec in (("fp32", A.PREC_FP32), ("fp64", A.PREC_FP64)):
            lanes, cand = ctx.autotune_knn(k, prec, sample_nq=min(chunk, 4_000_000))
            ctx.knn_dev(k, prec, nq=chunk)
            sync_all(ctx)
            t = time.perf_counter()
            kern = 0.0
            for q0 in range(0, n, chunk):
                ctx.knn_dev(k, prec, first=q0, nq=min(chunk, n - q0))
                kern += ctx.stage_times()["knn"]
            sync_all(ctx)
            wall = time.perf_counter() - t
            st = ctx.knn_dev(k, prec, nq=chunk, stats=True)  # counters in a run of their own (global atomics)
            ent[name] = {"queries_per_s": n / wall, "kernel_ms": kern, "lanes_per_query": lanes,
                         "candidate_ms": cand.tolist(), "points_tested_per_query": st["points_tested"] / chunk,
                         "rechecked_per_query": st["rechecked"] / chunk}
            log(f"[bench] knn k={k} {name}: {n / wall:.3e} queries/s over {n} queries in chunks of {chunk}, "
                f"kernel {kern:.1f} ms, lanes {lanes}")
            ksamp = np.linspace(0, chunk - 1, 400 if k <= 250 else 100).astype(np.int64)
            rows = [ctx.knn_fetch(int(q), 1) for q in ksamp]
            ent[name]["_sample"] = (ksamp, np.vstack([r[0] for r in rows]), np.vstack([r[1] for r in rows]))
        ent["fp32_vs_fp64_speedup"] = ent["fp64"]["kernel_ms"] / ent["fp32"]["kernel_ms"]
        knn[k] = ent
    qps = n / fp32["wall_s"]
    verify = {}
    if not args.no_verify:
        c32, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, radius=args.radius, nq=n)
        c64, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP64, radius=args.radius, nq=n)
        verify["fp32_eq_fp64_all_queries"] = bool(np.array_equal(c32, c64))
        verify["avg_result_size"] = int(c64.astype(np.int64).sum() // max(n, 1))

    # One GPU's share of the all-points search on 8 GPUs: the queries split into 8 contiguous SFC
    # ranges, each searched against the whole (replicated) cloud and tree, FP32 + recheck as tuned;
    # once with equal query counts and once balanced by work (equal neighbour counts + 64 per query,
    # from this run's all-points counts, as a previous pass would give them)
    shard = None
    if n >= 8:
        counts_w = c32 if not args.no_verify else ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32,
                                                             radius=args.radius, nq=n)[0]
        cum = np.cumsum(counts_w.astype(np.int64) + 64)
        splits = {"equal_queries": np.linspace(0, n, 9).astype(np.int64),
                  "balanced": np.concatenate([[0], np.searchsorted(cum, cum[-1] * np.arange(1, 8) / 8), [n]])}
        sidx = A.DeviceArray(np.arange(n, dtype=np.uint32))  # the sorted-cloud indices, in SFC order
        shard = {"gpus": 8}
        try:
            for sname, bounds in splits.items():
                sms, snb = [], 0
                for s in range(8):
                    q0, nq_s = int(bounds[s]), int(bounds[s + 1] - bounds[s])
                    for _ in range(2):
                        ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, args.radius, nq=nq_s, idx_ptr=sidx.ptr(q0))
                    sms.append(float(ctx.stage_times()["search"]))
                    st_s = ctx.radius_dev(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, args.radius, nq=nq_s, stats=True,
                                          idx_ptr=sidx.ptr(q0))
                    snb += int(st_s["total_neighbors"])
                shard[sname] = {"queries": [int(bounds[s + 1] - bounds[s]) for s in range(8)], "kernel_ms": sms,
                                "max_ms": max(sms), "sum_ms": sum(sms), "balance": float(np.mean(sms) / max(sms)),
                                "neighbors_eq_all_points": snb == int(fp32["stats"]["total_neighbors"])}
                log(f"[bench] 8 query shards, {sname}: kernel ms {[round(v, 2) for v in sms]}")
        finally:
            sidx.close()

    cpu = None
    competitor = None
    if not args.no_cpu or not args.no_verify:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        sx = np.ascontiguousarray(xyz[perm0.astype(np.int64)])
        t = time.time()
        tree = O.Tree(keys0, args.max_leaf, A.ENC_HILBERT, c, r)
        oracle_build_s = time.time() - t
        log(f"[bench] oracle tree built in {oracle_build_s:.1f} s (nTotal {tree.nTotal} vs gpu {info['n_total']})")
        verify["tree_size_eq_oracle"] = bool(tree.nTotal == info["n_total"] and tree.nLeaf == info["n_leaf"])
        if not args.no_verify:
            samp = np.arange(0, n, max(1, n // 2000), dtype=np.int64)[:2000].astype(np.uint32)
            ref, ref_sizes = O.radius(tree, sx, O.K_SPHERE, O.A_STRUCT, sx[samp.astype(np.int64)], args.radius)
            ok = True
            for prec in (A.PREC_FP32, A.PREC_FP64):
                cnt, lists, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, prec, idx=samp, radius=args.radius, lists=True)
                ok &= bool(np.array_equal(cnt, ref_sizes)) and all(np.array_equal(a, b) for a, b in zip(lists, ref))
            verify["lists_eq_oracle_sample"] = ok
            verify["sample_queries"] = int(len(samp))
            knn_ok, knn_err = True, 0.0
            for k, ent in knn.items():
                ksamp = ent["fp32"]["_sample"][0]
                t = time.time()
                oi, od = O.knn(tree, sx, sx[ksamp], k)  # the reference's best-first search
                ent["cpu_oracle_queries_per_s"] = len(ksamp) / max(time.time() - t, 1e-9)
                # distances against the best-first search on the whole sample; index lists against brute
                # force (order (d^2, index)) on its first 8 queries (the best-first order of ties is free)
                bi, _ = O.knn_brute(sx, sx[ksamp[:8]], k)
                err = 0.0
                for name in ("fp32", "fp64"):
                    _, gi, gf = ent[name]["_sample"]
                    knn_ok &= bool(np.array_equal(gf, od.astype(np.float32)))
                    knn_ok &= bool(np.array_equal(gi[:8].astype(np.int64), bi))
                    err = max(err, float(np.max(np.abs(gf.astype(np.float64) - od))))
                ent["max_abs_d2_error_fp32_output"] = err
                ent["sample_queries"] = int(len(ksamp))
                knn_err = max(knn_err, err)
            if knn:
                verify["knn_eq_oracle_sample"] = knn_ok
        if not args.no_cpu:
            threads = int(O.lib().orc_max_threads())
            reps = max(2, args.cpu_reps)
            # the reference's CPU path restated (the oracle, OpenMP on every host core): encode + sort
            # (enc_build_benchmarks.hpp:183-227) and the octree build (:300-345) on the same cloud, one
            # warm-up then the median of `reps` timed runs per stage
            cpu_stages = {}
            cpu_stages["bbox"], (cc, rr) = median_of(lambda: O.mbb(xyz), reps)
            cpu_stages["encode"], ckeys = median_of(lambda: O.encode(O.ENC_HILBERT, xyz, cc, rr), reps)
            cpu_stages["sort"], (csk, cperm) = median_of(lambda: O.sort(ckeys, nthreads=threads), reps)
            cpu_stages["gather"], _ = median_of(lambda: O.gather3(xyz, cperm), reps)
            del ckeys
            verify["cpu_keys_eq_gpu"] = bool(np.array_equal(csk, keys0) and np.array_equal(cperm, perm0))
            del csk, cperm
            cpu_stages["build"], _ = median_of(lambda: O.Tree(keys0, args.max_leaf, A.ENC_HILBERT, c, r), reps)
            log(f"[bench] cpu stages (s) {cpu_stages}")
            # the search: a strided sample of the all-points queries sized for ~cpu_seconds per run, the
            # reference harness loop (neighbor_benchmarks.hpp:507-561); the calibration run is the warm-up
            cal = np.arange(0, n, max(1, n // 20000), dtype=np.int64)
            t = time.time()
            O.radius_harness(tree, sx, O.K_SPHERE, O.A_STRUCT, sx[cal], args.radius)
            per_q = max((time.time() - t) / len(cal), 1e-9)
            nq_cpu = int(min(n, max(len(cal), args.cpu_seconds / reps / per_q)))
            stride = max(1, n // nq_cpu)
            samp = np.arange(0, n, stride, dtype=np.int64)
            centres = sx[samp]
            ts = []
            for _ in range(reps):
                t = time.perf_counter()
                O.radius_harness(tree, sx, O.K_SPHERE, O.A_STRUCT, centres, args.radius)
                ts.append(time.perf_counter() - t)
            dt = float(np.median(ts))
            cpu = {"value": len(samp) / dt, "unit": "queries/s", "cores": threads, "kind": "port",
                   "cpu_model": cpu_model(), "repeats": reps, "repeat_s": ts,
                   "sample": f"{len(samp)} of the {n} all-points queries (every {stride}th SFC index), "
                             f"neighborsStruct sphere r={args.radius} with the reference harness loop "
                             f"(build each range list, keep its size), OpenMP dynamic schedule; median of {reps} "
                             f"timed runs after a warm-up",
                   "stages_s": cpu_stages,
                   "encode_sort_s": cpu_stages["bbox"] + cpu_stages["encode"] + cpu_stages["sort"] + cpu_stages["gather"],
                   "build_s": cpu_stages["build"],
                   "compile_flags": "gcc -O2 -mavx2 -mbmi2 -fopenmp -ffp-contract=off (the reference's Release flags)"}
            log(f"[bench] cpu baseline {cpu}")
        competitors = []
        if not args.no_competitor:
            # the reference's pointer-based octree (restated in the oracle): built by inserting the sorted
            # points one by one from the cloud box (serial, as the reference), searched with its
            # neighbors() (boxOverlap descent) per centre, OpenMP over the centres
            t = time.time()
            pt = O.PointerOctree(sx, args.max_leaf, c, r)
            pt_build = time.time() - t
            nsamp = np.arange(0, n, 25, dtype=np.int64)
            centres = sx[nsamp]
            pt_dt, ptc = median_of(lambda: pt.radius_counts(O.K_SPHERE, centres, args.radius), 2)
            gpu_c, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, idx=nsamp.astype(np.uint32), radius=args.radius)
            competitors.append({"name": "pointer octree (the reference's Octree, restated), CPU", "build_s": pt_build,
                                "build_threads": 1, "radius_queries_per_s": len(nsamp) / pt_dt,
                                "cores": int(O.lib().orc_max_threads()), "leaves": pt.n_leaf, "depth": pt.depth,
                                "sample": f"{len(nsamp)} centres (every 25th SFC index), sphere r={args.radius}, "
                                          f"median of 2 runs after a warm-up",
                                "counts_eq_gpu": bool(np.array_equal(ptc, gpu_c.astype(np.uint64)))})
            log(f"[bench] competitor {competitors[-1]}")
            del pt
        if not args.no_competitor and O.nanoflann_available():
            # the reference's competitor: its vendored nanoflann KD-tree (leaf size maxLeaf) over the sorted
            # cloud, radiusSearch per centre as its benchmark drives it, on a sample of the centres
            t = time.time()
            nf = O.NanoflannTree(sx, args.max_leaf)
            nf_build = time.time() - t
            nsamp = np.arange(0, n, 25, dtype=np.int64)
            centres = sx[nsamp]
            nf_dt, nfc = median_of(lambda: nf.radius_counts(centres, args.radius), 2)
            gpu_c, _ = ctx.radius(A.K_SPHERE, A.A_STRUCT, A.PREC_FP32, idx=nsamp.astype(np.uint32), radius=args.radius)
            competitors.append({"name": "nanoflann KD-tree (reference's vendored 1.7.1), CPU", "build_s": nf_build,
                                "radius_queries_per_s": len(nsamp) / nf_dt, "cores": int(O.lib().orc_max_threads()),
                                "sample": f"{len(nsamp)} centres (every 25th SFC index), sphere r={args.radius}, "
                                          f"median of 2 runs after a warm-up",
                                "counts_eq_gpu": bool(np.array_equal(nfc, gpu_c.astype(np.uint64)))})
            log(f"[bench] competitor {competitors[-1]}")
            del nf
        competitor = competitors or None

    # Roofline of the dominant kernel (the FP32 search): 9 flops (3 sub + 3 mul + 2 add + 1 compare) per
    # point test of the reference traversal (T, identical for both precisions), over the kernel's
    # device time (HIP events on the context's stream), against MI355X's FP32 compute roof; the kernel
    # runs on the vector ALUs. traffic: the kernel's HBM bytes per launch from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated on a copy kernel), used only when the
    # summary was taken from this kernel's current sources and tile shape.
    T = fp32["stats"]["points_tested"]
    kern_s = fp32["kernel_ms"] * 1e-3
    achieved = 9.0 * T / kern_s / 1e12
    traffic, traffic_note = None, "no counter summary"
    digest = source_digest()
    try:
        tr = json.load(open(os.path.join(ROOT, "profiles", "search_c2_50M_traffic.json")))
        cfg = f"k_count<1, 3, 1, {tuned['fp32']['queries_per_wave']}, false>"
        if n != 50_000_000 or las is not None:
            traffic_note = "counter summary is of the 50 M-point tile"
        elif not tr["kernel"].startswith(cfg):
            traffic_note = f"counter summary is of {tr['kernel']}, tuned shape is {cfg}"
        elif tr.get("source_digest") != digest:
            traffic_note = f"counter summary is of sources {tr.get('source_digest')}, these are {digest}"
        else:
            traffic, traffic_note = tr["hbm_bytes_per_launch"], f"rocprofv3 counters, sources {digest}"
    except (OSError, ValueError, KeyError):
        pass
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": traffic_note,
                "kernel": "radius count (FP32 + FP64 re-check)",
                "compute_unit": "vector ALUs",
                "work": f"9 flops x {T} point tests of the reference traversal"}

    es_ms = med(bbox_ms) + med(enc_ms) + med(sort_ms) + med(gather_ms)
    # algorithmic bytes of bbox + encode + sort + gather as the reference's algorithm moves them: 24 + 32 +
    # 8 (histogram) + P * 24 (key + index read and written per pass) + 52 (gather) B/pt with 8-bit digits,
    # P = 8 passes. The partition sort moves fewer: 24 + 32 + 20 + 24 (two passes) + 24 (LDS sorts) + 56
    # (gather, fused into them) B/pt, plus 32 B for keys in 16-bit segments that are split.
    es_bytes = (116 + 24 * 8) * n
    es_gbs = es_bytes / (es_ms * 1e-3) / 1e9
    copy_gbs, _ = ctx.copy_bandwidth(1 << 30)  # the practical streaming ceiling, measured in this run
    log(f"[bench] streaming copy ceiling {copy_gbs:.0f} GB/s")
    roofline_es = {"bound": "hbm", "achieved": es_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": es_gbs / PEAK_HBM_GBS,
                   "traffic": None,
                   "kernel": "bbox + encode + partition sort (2 device-wide passes on the top 16 key bits, "
                             "per-segment LDS sorts that also gather the sorted cloud)",
                   "work": f"{116 + 24 * 8} B/pt x {n} points (8-pass LSD sort + gather, the reference's algorithm)",
                   "path_work": "180 B/pt + 32 B per key in a split 16-bit segment", "copy_ceiling_gbs": copy_gbs,
                   "frac_of_copy_ceiling": es_gbs / copy_gbs}

    out = {
        "metric": "radius queries/s @50M pts, FP32+recheck vs FP64 GPU (exact sets); build ms",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": fp32["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-recheck",
        "data": "synthetic" if las is None else "LAS file",
        "config": {"workload": (f"synthetic airborne-LiDAR tile, {n} FP64 points over {extent:.0f}x{extent:.0f} m, "
                                if las is None else f"LAS file {las['file']}, {n} points, ") +
                               f"Hilbert 21 bits/axis, maxLeaf {args.max_leaf}, all-points sphere r={args.radius} m, "
                               f"neighborsStruct, per-query neighbour counts",
                   "points": n, "queries": n, "radius": args.radius, "max_points_leaf": args.max_leaf,
                   "encoder": "hilbert"},
        "fp64_queries_per_s": n / fp64["wall_s"],
        "fp32_vs_fp64_speedup": fp64["wall_s"] / fp32["wall_s"],
        "recheck_rate": fp32["stats"]["rechecked"] / max(1, fp32["stats"]["points_tested"]),
        "neighbors_per_s": fp32["stats"]["total_neighbors"] / fp32["wall_s"],
        "build_ms": med(build_ms),
        "encode_sort_ms": es_ms,
        "encode_sort_hbm_frac": es_bytes / (es_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
        "points_encoded_per_s": n / (med(enc_ms) * 1e-3),
        "tree": {k: info[k] for k in ("n_leaf", "n_internal", "n_total", "max_depth_seen")},
        "stages_ms": {"bbox": med(bbox_ms), "encode": med(enc_ms), "sort": med(sort_ms), "gather": med(gather_ms),
                      "build": med(build_ms), "search_fp32": fp32["kernel_ms"],
                      "search_fp64": fp64["kernel_ms"]},
        "verify": verify,
        "las": las,
        "autotune": tuned,
        "union_efficiency": {"fp32_test_slots_used": fp32["stats"]["points_tested"] / max(1, fp32["stats"]["test_slots"]),
                             "fp32_class_slots_used": fp32["stats"]["nodes_classified"] / max(1, fp32["stats"]["class_slots"])},
        "knn": {str(k): {kk: ({x: y for x, y in v.items() if x != "_sample"} if kk in ("fp32", "fp64") else v)
                         for kk, v in ent.items()} for k, ent in knn.items()},
        "roofline": roofline,
        "roofline_encode_sort": roofline_es,
        "shard_8": shard,
        "cpu_baseline": cpu,
        "competitor_cpu": competitor,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
