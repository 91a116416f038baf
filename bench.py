#!/usr/bin/env python3
"""Throughput of the DistMesh force loop on one MI355X.

Workload (default): the 3D many-primitive CSG of seismicmesh_amd.workloads
(48 balls and 16 boxes, minus 8 cylinders, plus 4 tori in [-1, 1]^3, seed
3) at uniform h = 0.015 (~0.5 M vertices).  Two paths run the same
iterations from the same initial points:

* fused: the tetrahedralisation built by 16 host threads inserting the
  points of 16 slabs into one shared tetrahedralisation, its cells sorted
  into pinned memory and copied to the device in pieces, and a neural SDF
  surrogate (3 -> 256 x4 -> 1, BF16 matrix-core inference) trained here by
  the library's own kernels on 4 M samples of the exact SDF deciding the
  centroid filter and the projection test away from the boundary (exact
  SDF inside the band);
* plain: the serial tetrahedralisation and the exact SDF everywhere.

A step is one force iteration over the whole point set: re-triangulation,
centroid filter, sorted unique edges, bar forces, gather, update and Newton
projection.  The timed region is `steps` iterations starting from the
initial points (the reference's loop performs max_iter - 1 = 49 of them);
initialisation (lattice, surrogate training) and termination are outside
it, as in the metric's definition.  `--workload` selects the other
workloads (2D BP2004-shaped, unit disk, 512 independent 2D domains,
EAGE-shaped 3D).

Output: one JSON line with the metric, the roofline of the dominant
streaming kernel and of every streaming kernel (HIP events on the loop's
own stream) and a CPU baseline (the numpy oracle on a bounded sample of the
same workload, one pinned core).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


# kernels timed alone by the handle's event profiler, by phase name
SINGLE_KERNEL = {"centroid_filter": "k_keep", "edge_geom": "k_edge_geom",
                 "edge_force": "k_edge_force", "vertex_update": "k_vertex_update",
                 "flip_mark": "k_mark", "unique_edges": "edge build (k_count, scan, "
                 "k_scatter, k_bucket_sorts, k_emit)"}


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if any."""
    fn = os.path.join(HERE, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    try:
        with open(fn) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def alg_bytes(phase, N, E, C, T, dim=2, grid_taps=True):
    """Algorithmic HBM bytes of one pass of a phase over the whole mesh
    (FP64 points, FP32 per-edge arrays and sizing grid, int32 ids): every
    datum that has to cross HBM is counted once -- a vertex's coordinates
    once however many edges or cells gather them, an edge's force vector
    once although both endpoints read it.  Sizing-grid taps count 2^dim
    FP32 values per edge (an upper bound on the distinct taps; none for a
    constant size)."""
    pt = 8 * dim
    taps = 4 * 2 ** dim if grid_taps else 0
    if phase == "edge_geom":
        # edge ids in, L and h out; points and sizing taps gathered
        return E * (8 + taps + 8) + N * pt
    if phase == "edge_force":
        # edge ids, L and h in, force vector out; points gathered
        return E * (8 + 8 + 4 * dim) + N * pt
    if phase == "vertex_update":
        # per vertex: two list offsets, point in and out; per edge: its force
        # vector and its entry in the second-endpoint list
        return N * (8 + 2 * pt) + E * (4 * dim + 4)
    if phase == "unique_edges":
        # kept cells in, sorted edges and first-endpoint offsets out
        return C * 4 * (dim + 1) + E * 8 + N * 4
    if phase == "centroid_filter":
        # cell ids in, keep flag out, points gathered
        return T * (4 * (dim + 1) + 1) + N * pt
    if phase == "flip_mark":
        # per triangle: ids 12, neighbours 12, three 64-bit candidates 24,
        # its 64-bit claim 8; points 16 per vertex.  Orientations are not
        # read when no triangle is inverted or uncertain
        return T * (12 + 12 + 24 + 8) + N * 16
    return 0


STREAMING = ("centroid_filter", "unique_edges", "edge_geom", "edge_force", "vertex_update")


# the same phases on the surrogate-steered path (one launch of each per
# iteration)
SURROGATE_KERNELS = {"centroid_filter": "k_mlp_eval_lds (centroids) + k_keep_surrogate",
                     "vertex_update": "k_vertex_update + k_mlp_eval_lds (points) + "
                                      "k_project_surrogate"}


def streaming_rooflines(prof, iters, N, E, C, T, dim, grid_taps, names=None):
    """Achieved GB/s of each streaming phase over the profiled iterations:
    algorithmic bytes per iteration / event-timed ms per iteration."""
    names = dict(SINGLE_KERNEL, **(names or {}))
    out = {}
    for ph in STREAMING:
        ms, launches = prof.get(ph, (0.0, 0))
        if ms <= 0 or iters <= 0:
            continue
        b = alg_bytes(ph, N, E, C, T, dim, grid_taps)
        per_it = ms / iters
        gbs = b / (per_it * 1e-3) / 1e9
        out[ph] = {"kernel": names.get(ph, ph), "alg_bytes_per_iteration": int(b),
                   "ms_per_iteration": round(per_it, 4),
                   "launches_per_iteration": round(launches / iters, 2),
                   "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_baseline(st, p0, iters, reps=3):
    """The oracle (numpy FP64 restatement + scipy Qhull) on `iters` force
    iterations of the same workload (or of the given subset p0 of its
    points): one thread pinned to one core, BLAS pools limited to one
    thread, median of `reps` runs."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from oracle import distmesh as odm
    from oracle.sizing import build_interpolant

    odom = _oracle_domain(st["dom"])
    fh = st["fh"]
    if fh.kind == "grid":
        rgi = build_interpolant(fh.grid.values.astype(np.float64), fh.grid.bbox)
        ofh = rgi
    else:
        val = fh.value
        ofh = lambda x: np.array([val] * len(x))  # noqa: E731
    dim = st["dim"]
    s = dict(fd=odom.eval, fh=ofh, h0=st["h0"], dim=dim, L0mult=1 + 0.4 / 2 ** (dim - 1),
             geps=st["geps"],
             deps=st["deps"], pfix=p0[:min(st["nfix"], len(p0))], levels=[odom.eval],
             delta_t=st["delta_t"])
    old_aff = os.sched_getaffinity(0)
    core = min(old_aff)
    os.sched_setaffinity(0, {core})
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    if ctx:
        ctx.__enter__()
    times = []
    try:
        for _ in range(reps):
            p = p0.copy()
            t0 = time.perf_counter()
            for _ in range(iters):
                t_all = odm.delaunay(p)
                p, _, _ = odm.force_step(p, t_all, s, list(range(min(st["nfix"], len(p)))))
            times.append(time.perf_counter() - t0)
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
        os.sched_setaffinity(0, old_aff)
    dt = float(np.median(times))
    return len(p0) * iters / dt, dt, core


def _oracle_domain(dom):
    """The oracle's copy of a product geometry object (Disk, Rectangle, Ball,
    Cube, Cylinder, Torus -- translated, not rotated or stretched -- and
    their unions / differences)."""
    from oracle import geometry as og
    name = type(dom).__name__
    tr = getattr(dom, "translation_vec", None)
    tr = None if tr is None else [float(v) for v in tr]
    if name == "Disk":
        return og.Disk([dom.xc, dom.yc], dom.r, translate=tr)
    if name == "Rectangle":
        return og.Rectangle(dom.bbox0, translate=tr)
    if name == "Ball":
        return og.Ball([dom.xc, dom.yc, dom.zc], dom.r, translate=tr)
    if name == "Cube":
        return og.Cube(dom.bbox0, translate=tr)
    if name == "Cylinder":
        return og.Cylinder(h=2.0 * dom.h[1], r=dom.h[0], translate=tr)
    if name == "Torus":
        return og.Torus(dom.t[0], dom.t[1], translate=tr)
    if name in ("Union", "Difference", "Intersection"):
        return getattr(og, name)([_oracle_domain(d) for d in dom.domains])
    raise ValueError("no oracle copy of %s" % name)


def cpu_baseline_many(sts, iters, reps=3, min_points=50000):
    """The oracle on the first domains of the many-domains workload (at
    least `min_points` points), `iters` iterations each, one pinned thread,
    median of `reps` runs."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from oracle import distmesh as odm
    cases = []
    total = 0
    for st in sts:
        od = _oracle_domain(st["dom"])
        h = st["h0"]
        s = dict(fd=od.eval, fh=(lambda x, h=h: np.array([h] * len(x))), h0=h, dim=2,
                 L0mult=1.2, geps=st["geps"], deps=st["deps"], pfix=st["p"][:st["nfix"]],
                 levels=[od.eval], delta_t=st["delta_t"])
        cases.append((st["p"], s, st["nfix"]))
        total += len(st["p"])
        if total >= min_points:
            break
    old_aff = os.sched_getaffinity(0)
    core = min(old_aff)
    os.sched_setaffinity(0, {core})
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    if ctx:
        ctx.__enter__()
    times = []
    try:
        for _ in range(reps):
            t0 = time.perf_counter()
            for p0, s, nfix in cases:
                p = p0.copy()
                for _ in range(iters):
                    p, _, _ = odm.force_step(p, odm.delaunay(p), s, list(range(nfix)))
            times.append(time.perf_counter() - t0)
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
        os.sched_setaffinity(0, old_aff)
    dt = float(np.median(times))
    return total * iters / dt, dt, core, total


MFMA_BF16_PEAK_TFLOPS = 2500.0
# the surrogate network 3 -> 256 -> 256 -> 256 -> 256 -> 1: FLOP per evaluation
MLP_FLOP_PER_POINT = 2 * (3 * 256 + 3 * 256 * 256 + 256)


def mlp_measure(npts=1 << 22, reps=10):
    """The neural SDF surrogate's inference kernel (k_mlp_eval_lds, BF16 matrix
    cores, 3 -> 256 x4 -> 1) on npts points with seeded random weights:
    average launch time from HIP events on the launch stream, and its
    FLOP rate against the dense BF16 peak."""
    import ctypes

    import torch
    from seismicmesh_amd import _lib
    from seismicmesh_amd.neural import NeuralSDF

    net = NeuralSDF(3, [-1.0, 1.0] * 3, hidden=3, seed=0, omega=8.0)
    x = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, (npts, 3))).cuda()
    out = torch.empty(npts, dtype=torch.float32, device="cuda")
    st = net.struct()
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def launch():
        _lib.check(_lib.lib.sm_mlp_eval(ctypes.byref(st), ctypes.c_void_p(x.data_ptr()), npts,
                                        ctypes.c_void_p(out.data_ptr()), sp), "sm_mlp_eval")

    for _ in range(3):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flop = float(MLP_FLOP_PER_POINT) * npts
    tf = flop / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "kernel": "k_mlp_eval_lds", "points": npts,
            "flop_per_point": int(flop / npts), "avg_launch_us": round(ms * 1e3, 2),
            "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}


def bench_3d(args):
    """The 3D workloads: the fused path (slab-parallel tetrahedralisation
    overlapped with the device filter, and for the CSG a trained surrogate)
    against the plain path (serial tetrahedralisation, exact SDF), same
    points, same iterations."""
    import torch
    from seismicmesh_amd import generation, workloads
    from seismicmesh_amd import mesh_utils as mu

    t_start = time.perf_counter()

    def progress(msg):
        print("[bench %.1fs] %s" % (time.perf_counter() - t_start, msg), file=sys.stderr,
              flush=True)

    if args.workload == "csg3d":
        domain, bbox3 = workloads.csg3d()
        ef = float(args.h)
        wl = {"workload": "csg3d_neural", "primitives": "48 balls U 16 boxes, minus 8 "
              "cylinders, U 4 tori in [-1, 1]^3 (seed 3)", "h": ef, "max_iter": 50}
        grid_taps = False
    else:
        domain, ef, info = workloads.eage3d(scale=args.scale)
        wl = {"workload": "eage_3d", "grid": list(info["grid"]), "bbox_m": list(info["bbox"]),
              "hmin_m": 150.0, "grade": 0.15, "max_iter": 50}
        grid_taps = True
    st = generation.prepare(domain, ef, verbose=0, precision=args.precision)
    p0 = st["p"]
    N = len(p0)
    progress("prepared %s: %d points" % (wl["workload"], N))
    threads = max(1, min(args.slabs, os.cpu_count() or 1))
    fused = generation.make_mesher(st, slabs=1)
    axis = int(st["opts"]["axis"])
    fused.set_slabs(args.slabs, axis=axis, threads=threads)
    plain = generation.make_mesher(st, slabs=1)
    # workgroup sizes tuned on this workload (two iterations in), shared
    fused.set_points(p0)
    fused.run(2)
    tuned = fused.autotune(reps=10)
    sizes = {k: v[0] for k, v in tuned.items()}
    plain.set_workgroups(sizes)
    progress("autotune: %s" % sizes)

    neural = None
    if args.workload == "csg3d":
        from seismicmesh_amd.neural import NeuralSDF
        t0 = time.perf_counter()
        net = NeuralSDF(3, bbox3, hidden=3, seed=0, omega=8.0)
        # trained on the SDF truncated at 0.1: the network only has to be
        # accurate near the surface
        net.train(domain, nsamples=args.train_samples, steps=args.train_steps, batch=8192,
                  lr=2e-3, lr_final=1e-4, eik=0.1, near=0.05, seed=0, trunc=0.1)
        t_train = time.perf_counter() - t0
        # a band per region of a 128^3 grid: twice the largest error measured
        # there (and in the neighbouring regions), at least 2 h.  Fine
        # regions keep a bad spot near one of the 76 primitives from closing
        # a large block to the network: at 16^3 only ~1 % of the interior
        # points were decided without the exact program, at 128^3 ~38 %
        # (held-out points, no wrong decision)
        bands, blo, bhi, err = net.calibrate_bands(domain, 2.0 * ef, res=128, nval=32000000)
        finite = bands[bands < 1e29]
        if finite.size == 0:
            finite = np.array([np.inf])
        exact_only = float(np.mean(bands >= 1e29))
        progress("surrogate trained in %.1f s: max error %.4f, bands %.4f .. %.4f (median %.4f), "
                 "%.0f %% of the regions exact only"
                 % (t_train, err, finite.min(), finite.max(), float(np.median(finite)),
                    100 * exact_only))
        fused.set_surrogate(net, 2.0 * ef, bands=(bands, blo, bhi))
        neural = {"training_s": round(t_train, 1), "steps": args.train_steps,
                  "samples": args.train_samples, "near": 0.05, "trunc": 0.1,
                  "max_error": round(err, 4),
                  "band_min": 2.0 * ef, "bands": "128^3 regions, 2 x the largest error measured "
                  "on 32 M validation points (half within 2 h of the surface) in each region "
                  "and its neighbours; a region whose band reaches the truncation is left to "
                  "the exact program",
                  "band_median": round(float(np.median(finite)), 4),
                  "band_max": round(float(finite.max()), 4),
                  "regions_exact_only": round(exact_only, 4)}

    def min_dihedral(h):
        """Smallest dihedral angle (degrees) of the loop's kept cells, and
        the 1st percentile of the cells' smallest angles.  The angle formula
        (the reference's acos) gives NaN for cells flat to rounding, counted
        as 0 degrees."""
        h.triangulate()
        p, t = h.points(), h.cells()
        a = np.degrees(np.nan_to_num(mu.dihedral_angles(p, t), nan=0.0)).min(1)
        return round(float(a.min()), 4), round(float(np.percentile(a, 1)), 4), len(t)

    def timed(h, name):
        h.set_points(p0)
        if args.warmup > 0:
            h.run(args.warmup)
        h.set_points(p0)
        s0 = h.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.run(args.steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s1 = h.stats()
        progress("%s loop: %d iterations in %.3f s" % (name, args.steps, dt))
        d = {k: s1[k] - s0[k] for k in s1}
        d["min_quality"], d["p1_quality"], d["cells"] = min_dihedral(h)
        return dt, d

    elapsed, stats = timed(fused, "fused")
    p_fused = fused.points()
    if neural is not None:
        sg = fused.surrogate_stats()
        neural["exact_evaluations_skipped"] = round(sg["skipped"] / max(sg["skipped"] +
                                                                       sg["exact"], 1), 4)
        neural["max_error_in_band"] = round(sg["max_err"], 4)
    plain_elapsed, plain_stats = timed(plain, "plain")
    identical = bool(np.array_equal(p_fused, plain.points()))
    n, C, E = fused.sizes()
    value = N * args.steps / elapsed
    plain_value = N * args.steps / plain_elapsed

    # the fused region again with per-phase HIP events on the loop's stream
    fused.set_points(p0)
    fused.profile(True)
    s0 = fused.stats()
    fused.run(args.steps)
    prof = fused.profile_read()
    s1 = fused.stats()
    fused.profile(False)
    progress("event-profiled loop done")
    streaming = streaming_rooflines(prof, args.steps, N, E, C, C, 3, grid_taps,
                                    SURROGATE_KERNELS if neural else None)
    best = max(streaming, key=lambda k: streaming[k]["ms_per_iteration"])
    b = streaming[best]
    roofline = {"bound": "hbm", "kernel": b["kernel"], "achieved": b["achieved_GBps"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b["frac"],
                "traffic": load_traffic(SINGLE_KERNEL.get(best, best), wl["workload"]),
                "alg_bytes_per_iteration": b["alg_bytes_per_iteration"],
                "ms_per_iteration": b["ms_per_iteration"],
                "launches_per_iteration": b["launches_per_iteration"],
                "streaming": streaming,
                "phases_ms": {k: round(v[0], 3) for k, v in prof.items() if v[0] > 0}}
    if neural is not None and prof["centroid_network"][1] > 0:
        # the dominant kernel of the fused iteration is the network at the
        # centroids (one launch per iteration over every finite tetrahedron
        # of the build): its matrix-core roofline, timed by the same events
        ms_net, n_net = prof["centroid_network"]
        tets = s1["tets_last_build"]
        flops = tets * MLP_FLOP_PER_POINT
        tfs = flops / (ms_net / n_net * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": "k_mlp_eval_lds (centroids)",
                    "achieved": round(tfs, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tfs / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
                    "flop_per_launch": flops, "ms_per_launch": round(ms_net / n_net, 4),
                    "launches_per_iteration": round(n_net / args.steps, 2),
                    "note": "the network at every centroid is the largest kernel of the fused "
                            "iteration; the centroid filter around it (network + exact program "
                            "where the network does not decide) and the streaming kernels are "
                            "in streaming",
                    "streaming": streaming,
                    "phases_ms": {k: round(v[0], 3) for k, v in prof.items() if v[0] > 0}}
    # (flip_mark and centroid_network are timed inside flip_round and
    # centroid_filter)
    dev_ms = sum(v[0] for k, v in prof.items()
                 if k not in ("flip_mark", "centroid_network")) / args.steps

    cpu = None
    if not args.no_cpu:
        # a bounded sample: the fixed points and the points of the lowest
        # quarter along the slab axis, one oracle iteration
        nf = st["nfix"]
        rest = p0[nf:]
        cut = np.quantile(rest[:, axis], 0.25)
        sample = np.vstack([p0[:nf], rest[rest[:, axis] <= cut]])
        rate, secs, core = cpu_baseline(st, sample, 1, reps=1)
        model, ncpu = cpu_info()
        cpu = {"value": round(rate, 1), "unit": "vertex-updates/s", "cores": 1, "kind": "port",
               "sample": "1 force iteration of the workload's first quarter along axis %d "
                         "(%d of its %d points) with the numpy oracle + scipy Qhull (stand-in "
                         "for CGAL), one thread pinned to core %d of %d (%s), %.1f s"
                         % (axis, len(sample), N, core, ncpu, model, secs)}
        progress("cpu baseline done")
    mlp = None if args.no_mlp else mlp_measure()

    out = {
        "metric": "vertex-updates/s on 1 MI355X, fused+neural vs plain GPU at equal min quality",
        "value": round(value, 1),
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": dict(
            wl, N=N, edges=E, cells=C,
            path=("tetrahedralisation by %d host threads inserting the points of %d slabs "
                  "(axis %d) into one shared tetrahedralisation, its cells sorted into pinned "
                  "memory and copied to the device piece by piece; device filter, edges, "
                  "forces, update, projection%s"
                  % (threads, args.slabs, axis,
                     "; surrogate-steered filter and projection (exact SDF in the band)"
                     if neural else "")),
            plain_value=round(plain_value, 1),
            speedup_vs_plain=round(value / plain_value, 3),
            plain_path="serial tetrahedralisation, exact SDF, same device kernels",
            points_identical_to_plain=identical,
            min_quality=stats["min_quality"], plain_min_quality=plain_stats["min_quality"],
            p1_quality=stats["p1_quality"], plain_p1_quality=plain_stats["p1_quality"],
            quality_measure="minimum dihedral angle (degrees) of the raw loop mesh after the "
                            "timed iterations (cells flat to rounding count as 0), and the 1st "
                            "percentile of the cells' minimum angles",
            host_tet_ms_per_iteration=round(stats["host_tet_us"] / 1e3 / args.steps, 2),
            plain_host_tet_ms_per_iteration=round(plain_stats["host_tet_us"] / 1e3 /
                                                  args.steps, 2),
            device_ms_per_iteration=round(dev_ms, 3),
            parallel_builds=stats["parallel_builds"],
            retried_points_per_iteration=round(stats["retried_points"] / args.steps, 1),
            workgroup_sizes=sizes),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mlp_roofline": mlp,
    }
    if neural is not None:
        out["config"]["surrogate"] = neural
    fused.close()
    plain.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="csg3d",
                    choices=["csg3d", "eage", "bp2004", "disk", "many"])
    ap.add_argument("--h", type=float, default=0.015,
                    help="uniform size of the 3D CSG workload")
    ap.add_argument("--slabs", type=int, default=16,
                    help="3D: slabs (and host threads) of the fused path's tetrahedralisation")
    ap.add_argument("--train-samples", type=int, default=4 << 20)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--domains", type=int, default=512,
                    help="number of domains of the many-domains workload")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64])
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--far-square", action="store_true",
                    help="many: triangulate each domain inside a far square, not its hull")
    ap.add_argument("--no-mlp", action="store_true",
                    help="skip the surrogate inference measurement")
    ap.add_argument("--workgroups", default=None, metavar="N,N,N,N,N,N",
                    help="fixed workgroup sizes (filter, force, update, orient, mark, apply) "
                         "instead of the autotuner's choice, e.g. to profile a tuned run")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (a kernel of known "
                         "traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401  (device runtime)
    from seismicmesh_amd import generation, workloads

    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.workload in ("csg3d", "eage"):
        print(json.dumps(bench_3d(args)))
        return

    sts = None
    if args.workload == "many":
        from seismicmesh_amd import batch
        if args.precision != 32:
            sys.exit("the many-domains workload runs FP32 forces")
        doms, hs = workloads.many_domains(n=args.domains)
        sts = batch.prepare_batch(doms, hs)
        p0 = np.vstack([s_["p"] for s_ in sts])
        m = batch.make_batch_mesher(sts, far_square=args.far_square)
        st = None
        wl = {"workload": "many_domains_2d", "domains": args.domains, "seed0": 1000,
              "n_per_domain": "log-uniform [500, 2e5]", "max_iter": 50,
              "triangulated_region": "far square" if args.far_square else "convex hull"}
    else:
        if args.workload == "bp2004":
            domain, ef, info = workloads.bp2004(scale=args.scale)
            wl = {"workload": "bp2004_2d", "grid": list(info["grid"]),
                  "bbox_m": list(info["bbox"]), "hmin_m": 75.0, "grade": 0.15, "max_iter": 50}
        else:
            domain, ef, info = workloads.disk()
            wl = {"workload": "unit_disk_h0.02", "max_iter": 50}
        st = generation.prepare(domain, ef, verbose=0, precision=args.precision)
        p0 = st["p"]
        m = generation.make_mesher(st)
    N = len(p0)
    t_start = time.perf_counter()

    def progress(msg):
        print("[bench %.1fs] %s" % (time.perf_counter() - t_start, msg), file=sys.stderr,
              flush=True)

    progress("prepared %s: %d points" % (wl["workload"], N))

    def timed(graph, h=None):
        h = m if h is None else h
        h.set_graph(graph)
        # warm-up: a full loop from the initial points (allocations, graph)
        h.set_points(p0)
        if args.warmup > 0:
            h.run(args.warmup)
        # timed: `steps` iterations from the initial points
        h.set_points(p0)
        s0 = h.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.run(args.steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s1 = h.stats()
        progress("%s loop: %d iterations in %.3f s" % ("graph" if graph else "plain",
                                                      args.steps, dt))
        d = {k: s1[k] - s0[k] for k in s1}
        d["min_quality"] = min_quality(h)
        return dt, d

    def min_quality(h):
        """Smallest cell quality of the loop's mesh after the timed
        iterations (outside the timed region): 2D radius ratio, 3D minimum
        dihedral angle in degrees."""
        from seismicmesh_amd import mesh_utils as mu
        h.triangulate()
        p, t = h.points(), h.cells()
        if sts is not None:
            return round(min(float(mu.simp_qual(q, c).min()) for q, c in m.split(p, t)
                             if len(c)), 4)
        if m.dim == 2:
            return round(float(mu.simp_qual(p, t).min()), 4)
        return round(float(np.degrees(mu.dihedral_angles(p, t)).min()), 3)

    # the autotuner picks each tunable kernel's workgroup size on this
    # workload (two iterations in, so edges and forces exist)
    m.set_points(p0)
    m.run(2)
    if args.workgroups:
        sizes = [int(v) for v in args.workgroups.split(",")]
        m.set_workgroups(sizes)
        tuned = {k: (v, None) for k, v in zip(m.TUNABLE, sizes)}
    else:
        tuned = m.autotune(reps=20)
    dim = m.dim
    progress("autotune: %s" % {k: v[0] for k, v in tuned.items()})

    plain_elapsed, plain_stats = timed(False)
    elapsed, stats = timed(True)
    n, C, E = m.sizes()
    value = N * args.steps / elapsed

    # same region again with per-phase HIP events on the loop's stream
    m.set_points(p0)
    m.profile(True)
    m.run(args.steps)
    prof = m.profile_read()
    m.profile(False)
    progress("event-profiled loop done")
    T = 2 * N if dim == 2 else C
    best = max(SINGLE_KERNEL, key=lambda k: prof[k][0])
    ms_tot, launches = prof[best]
    # the dominant kernel alone: back-to-back launches on the final state
    # (HIP events on the loop's stream), so that launch gaps do not count
    reps = 200
    if best == "flip_mark":
        avg_ms = m.time_kernel("flip_mark", reps)
        launches = reps
    else:
        avg_ms = ms_tot / max(launches, 1)
    traffic_alg = alg_bytes(best, N, E, C, T, dim)
    achieved = traffic_alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    kernel = SINGLE_KERNEL[best]
    others = sum(v[0] for k, v in prof.items() if k != "flip_mark")
    roofline = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": load_traffic(kernel, wl["workload"]),
                "avg_launch_us": round(avg_ms * 1e3, 3), "launches": launches,
                "alg_bytes_per_launch": traffic_alg,
                "phases_ms": {k: round(v[0], 3) for k, v in prof.items()},
                "flips_per_s": round(stats["flips"] / elapsed, 1),
                "flip_share": round(prof["flip_round"][0] / max(others, 1e-9), 4)}

    cpu = None
    if not args.no_cpu and sts is not None:
        rate, secs, core, nsamp = cpu_baseline_many(sts, args.cpu_iters)
        model, ncpu = cpu_info()
        cpu = {"value": round(rate, 1), "unit": "vertex-updates/s", "cores": 1, "kind": "port",
               "sample": "%d force iterations of the first domains of the workload (%d points) "
                         "with the numpy oracle + scipy Qhull (stand-in for CGAL), one thread "
                         "pinned to core %d of %d (%s), median of 3 runs of %.1f s"
                         % (args.cpu_iters, nsamp, core, ncpu, model, secs)}
    elif not args.no_cpu:
        iters, reps = (args.cpu_iters, 3) if dim == 2 else (1, 1)
        rate, secs, core = cpu_baseline(st, p0, iters, reps=reps)
        model, ncpu = cpu_info()
        cpu = {"value": round(rate, 1), "unit": "vertex-updates/s", "cores": 1, "kind": "port",
               "sample": "%d force iteration(s) of the same workload (N=%d) with the numpy "
                         "oracle + scipy Qhull (stand-in for CGAL), one thread pinned to core "
                         "%d of %d (%s), median of %d run(s) of %.1f s"
                         % (iters, N, core, ncpu, model, reps, secs)}

    if cpu is not None:
        progress("cpu baseline done")
    mlp = None if args.no_mlp else mlp_measure()

    out = {
        "metric": "vertex-updates/s on 1 MI355X, fused+neural vs plain GPU at equal min quality",
        "value": round(value, 1),
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": dict(wl, N=N, edges=E, cells=C,
                       path=("device flips; one captured graph launch per iteration"
                             if dim == 2 else "host tetrahedralisation, device filter, edges, "
                             "forces, update and projection"),
                       plain_value=round(N * args.steps / plain_elapsed, 1),
                       plain_path="same kernels launched one by one, host sync per stage",
                       host_rebuilds=stats["host_rebuilds"], flip_rounds=stats["flip_rounds"],
                       rollbacks=stats["rollbacks"], kinetic_repairs=stats["kinetic_repairs"],
                       workgroup_sizes={k: v[0] for k, v in tuned.items()},
                       min_quality=stats["min_quality"],
                       plain_min_quality=plain_stats["min_quality"],
                       quality_measure=("radius ratio (2D simp_qual)" if dim == 2
                                        else "minimum dihedral angle, degrees")),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mlp_roofline": mlp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
