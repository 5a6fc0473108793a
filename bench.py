"""Lloyd-iteration throughput on one MI355X: the whole metric in one JSON line.

Headline (value): N=2^27 points, d=4 float32 features (2 GiB, feature-major),
k=10, Gaussian blobs sigma=1 with centres uniform in [-50,50]^4, seed 0,
init = first 10 points; one step = one Lloyd iteration of the single_stage
pipeline (fused label + mass + exact-sum pass, FP64 divide, swap) with the
bounded fused strategy.  Points are resident in HBM before timing starts;
the data set's sum grid is pinned once beforehand (setup, like the upload).

The same line carries, as sub-objects with their own roofline and CPU
baseline (skip them with --no-extra, or run one alone with --workload):
  three_stage  the same 2 GiB workload through the three_stage pipeline;
  buffered     single_stage_buffered: the 2 GiB streamed from pinned host
               memory in 256 MiB chunks every iteration (host-link bound);
  c1           N=2^20, 30 iterations in full, single_stage and three_stage,
               with the CPU restatements of both run in full beside them;
  dense        N=2^24, d=64, k=1024: MFMA labeling with the exact refinement;
  batch        4096 uneven jobs, 25 iterations: the batch engine against the
               plain per-job loop on the same GPU;
  serve        Poisson arrivals of the batched mix plus 2% large jobs
               (N=2^24, d=32, k=256), continuous batching against the plain
               per-job loop, p50/p99, and a latency-versus-rate sweep.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cl-kmeans_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
MFMA16_PEAK_TFLOPS = 2500.0  # dense FP16 / BF16
FP32_PEAK_TFLOPS = 157.3     # FP32 vector
METRIC = "Lloyd iter/s+%HBM (4f/10c 2GiB); served/batched vs plain solves/s @tol; p99"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the CPU baselines: the process's CPU share (the
    OMP_NUM_THREADS the box sets, else the CPUs this process may run on)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_info():
    return {"host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def kernel_times(ctx):
    by = {}
    for name, _, ms in ctx.records():
        by.setdefault(name, []).append(ms)
    return by


def centroid_err(got, ref64, pts, labels):
    """max |got - ref| / max(|ref|, the cluster's mean |x| in that feature):
    a relative error that does not blow up for coordinates near 0 (the scale
    the tests use)."""
    d, k = ref64.shape
    cnt = np.bincount(labels, minlength=k).astype(np.float64)
    absx = np.stack([np.bincount(labels, weights=np.abs(pts[f].astype(np.float64)), minlength=k)
                     for f in range(d)])
    scale = np.maximum(np.abs(ref64), absx / np.maximum(cnt, 1))
    live = cnt > 0
    if not live.any():
        return 0.0
    return float((np.abs(got.astype(np.float64) - ref64)[:, live] / np.maximum(scale[:, live], 1e-30)).max())


def hot_kernel(by, names):
    for nm in names:
        if nm in by:
            return nm, float(np.mean(by[nm])), by[nm]
    return None, float("nan"), []


# --------------------------------------------------------------------------
# the 2 GiB workload (C2): single_stage, three_stage, buffered


class C2:
    """The 2^27-point data set, resident once, shared by the C2 legs."""

    def __init__(self, args):
        import clkmeans
        from clkmeans import hip
        self.clkmeans, self.hip = clkmeans, hip
        self.n, self.d, self.k = 1 << args.log2n, args.d, args.k
        self.threads = cpu_threads()
        t0 = time.time()
        self.pts, _ = clkmeans.generate_blobs(self.n, self.d, self.k, sigma=1.0, box=50.0, seed=0,
                                              threads=self.threads)
        log("generated %d x %d points in %.1fs" % (self.n, self.d, time.time() - t0))
        self.init = np.ascontiguousarray(self.pts[:, :self.k])  # first_x
        self.ctx = clkmeans.Context(0)
        self.stream = hip.Stream()
        self.dp = hip.DeviceArray.from_host(self.pts)
        n, d, k = self.n, self.d, self.k
        self.cen = [hip.DeviceArray.from_host(self.init), hip.DeviceArray((d, k), np.float32)]
        self.sums = hip.DeviceArray((d, k), np.float64)
        self.masses = hip.DeviceArray((k,), np.uint32)
        self.labels = hip.DeviceArray((n,), np.uint32)
        # the data set's sum grid, pinned once (setup; the strategies do the same)
        self.ctx.sum_grid_from_points(d, n, self.dp, stream=self.stream.handle)
        self.stream.synchronize()

    def close(self):
        self.ctx.sum_grid_clear()
        self.ctx.close()
        del self.dp, self.cen, self.sums, self.masses, self.labels

    def measure(self, pipeline, strategy, steps, warmup, stage_labeling="unroll_vector_bounded"):
        """W warm-up iterations, then K timed iterations from the start
        (first-x centroids, zeroed labels) with every kernel bracketed by HIP
        events on the stream it runs on; one more iteration for the checks."""
        clkmeans, hip = self.clkmeans, self.hip
        ctx, s, n, d, k = self.ctx, self.stream.handle, self.n, self.d, self.k
        fcfg = ctx.config(clkmeans.STAGE_FUSED, strategy)
        lcfg = ctx.config(clkmeans.STAGE_LABELING, stage_labeling)
        mcfg = ctx.config(clkmeans.STAGE_MASS_UPDATE, "part_private")
        ccfg = ctx.config(clkmeans.STAGE_CENTROID_UPDATE, "cluster_merge")
        cen = self.cen

        def iteration():
            old, new = cen[0], cen[1]
            if pipeline == "single_stage":
                ctx.lloyd_step(fcfg, d, n, k, self.dp, old, new, self.sums, self.labels, self.masses, stream=s)
            else:
                ctx.fill_zero(self.masses, stream=s)
                ctx.fill_zero(self.sums, stream=s)
                ctx.labeling(lcfg, d, n, k, self.dp, old, self.labels, stream=s)
                ctx.mass_update(mcfg, n, k, self.labels, self.masses, stream=s)
                ctx.centroid_update(ccfg, d, n, k, self.dp, self.sums, self.labels, self.masses, stream=s)
                ctx.divide(d, k, self.sums, self.masses, new, stream=s)
            cen[0], cen[1] = new, old

        def restart():
            cen[0].copy_from(self.init)
            ctx.fill_zero(self.labels)
            hip.synchronize()

        restart()
        for _ in range(warmup):
            iteration()
        self.stream.synchronize()
        # The step time: HIP events around the K iterations only (per-kernel
        # events between the launches cost 3-7% of a 0.45 ms step).
        restart()
        e0, e1 = hip.Event(), hip.Event()
        w0 = time.perf_counter()
        e0.record(self.stream)
        for it in range(steps):
            ctx.set_iteration(it)
            iteration()
        e1.record(self.stream)
        e1.synchronize()
        hip.synchronize()
        wall = time.perf_counter() - w0
        ms_step = e0.elapsed_ms(e1) / steps
        # The kernels' durations: the same K iterations from the same start,
        # every kernel bracketed by HIP events on its stream.
        restart()
        ctx.clear_records()
        ctx.profiling(True)
        p0, p1 = hip.Event(), hip.Event()
        p0.record(self.stream)
        for it in range(steps):
            ctx.set_iteration(it)
            iteration()
        p1.record(self.stream)
        p1.synchronize()
        hip.synchronize()
        ctx.profiling(False)
        by = kernel_times(ctx)
        c_in = cen[0].to_host()
        iteration()
        self.stream.synchronize()
        return {"ms_per_step": ms_step, "ms_per_step_instrumented": p0.elapsed_ms(p1) / steps, "wall": wall,
                "by_name": by, "c_in": c_in, "labels": self.labels.to_host(), "masses": self.masses.to_host(),
                "centroids": cen[0].to_host()}

    def full_parity(self, r, O):
        """Every label of the check iteration against the oracle's scan
        (multi-threaded), masses against the label histogram, centroids
        bit-exact against the oracle's grid sums and their error against
        FP64 sums."""
        _, ref_lab, _ = O.lloyd_mt(self.pts, r["c_in"], 1, self.threads)
        lab, m = r["labels"], r["masses"]
        ref = O.sums_grid(self.pts, lab, self.k)
        with np.errstate(invalid="ignore", divide="ignore"):
            ref64 = O.sums_f64(self.pts, lab, self.k) / m.astype(np.float64)
        err = centroid_err(r["centroids"], ref64, self.pts, lab)
        return {"labels_checked": int(self.n), "labels_equal": bool(np.array_equal(lab, ref_lab)),
                "masses_equal_bincount": bool(np.array_equal(m, np.bincount(lab, minlength=self.k))),
                "centroids_equal_oracle_grid_sums": bool(np.array_equal(r["centroids"],
                                                                        O.divide_from_f64(ref, m))),
                "centroid_max_rel_err_vs_fp64": err,
                "centroid_err_scale": "max(|c|, the cluster's mean |x|) per coordinate"}


def fused_headline(args, c2):
    compare = args.strategy != "cluster_merge" and not args.no_compare
    exhaustive = c2.measure("single_stage", "cluster_merge", args.steps, args.warmup) if compare else None
    run = c2.measure("single_stage", args.strategy, args.steps, args.warmup)
    names = {"feature_sum": ["FusedFeatureSum"], "cluster_merge_bounded": ["FusedClusterMergeBounded"]}
    hot, hot_ms, hot_each = hot_kernel(run["by_name"], names.get(args.strategy, ["FusedClusterMerge"]))
    n, d, k = c2.n, c2.d, c2.k
    bytes_per_launch = n * (4 * d + 4)
    achieved = bytes_per_launch / (hot_ms * 1e-3) / 1e9
    log("kernels (avg ms): " + ", ".join("%s=%.4f x%d" % (kk, np.mean(v), len(v))
                                        for kk, v in run["by_name"].items()))

    cpu, parity = None, None
    if not args.no_cpu:
        # The CPU baseline leg: the oracle timed on the host cores, and as the
        # checker of the GPU's results.
        from oracle import oracle as O
        parity = c2.full_parity(run, O)
        iters = max(1, args.cpu_iters)
        t = time.perf_counter()
        O.lloyd_mt(c2.pts, c2.init, iters, c2.threads)
        dt = time.perf_counter() - t
        cpu = {"value": iters / dt, "unit": "iter/s", "cores": c2.threads, "kind": "port",
               "sample": "%d full iterations of the same %d-point workload, fused cluster_merge restatement "
                         "(oracle/kmeans_oracle.c ora_lloyd_mt_f32), %d threads" % (iters, n, c2.threads),
               **host_info()}
    traffic = None
    pmc = os.path.join(REPO, "profiles", "fused_%s_n2p27_d4_k10_pmc.json" % args.strategy)
    if os.path.exists(pmc) and n == (1 << 27) and d == 4 and k == 10:
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    def summary(r, strategy):
        h, ms, _ = hot_kernel(r["by_name"], ["FusedClusterMerge"])
        ach = bytes_per_launch / (ms * 1e-3) / 1e9
        return {"strategy": strategy, "value": 1000.0 / r["ms_per_step"], "unit": "iter/s",
                "ms_per_step": r["ms_per_step"], "kernel": h, "kernel_ms": ms,
                "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS}

    return {
        "metric": METRIC,
        "value": 1000.0 / run["ms_per_step"],
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain fused Lloyd, N=2^%d d=%d k=%d float32 feature-major, sigma=1 blobs "
                               "in [-50,50]^%d, seed 0, first-x init" % (args.log2n, d, k, d),
                   "pipeline": "single_stage", "strategy": args.strategy, "n": n, "d": d, "k": k},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": hot,
                     "kernel_ms": hot_ms, "bytes_per_launch": bytes_per_launch,
                     "kernel_ms_iterations_1_to_3": [round(x, 4) for x in hot_each[:3]],
                     "kernel_ms_last": round(hot_each[-1], 4) if hot_each else None,
                     "kernel_ms_source": "a second pass over the same %d iterations from the same start, every "
                                         "kernel bracketed by HIP events (%.4f ms per step with those events)"
                                         % (args.steps, run["ms_per_step_instrumented"])},
        "exhaustive": summary(exhaustive, "cluster_merge") if exhaustive else None,
        "cpu_baseline": cpu,
        "wall_s": run["wall"],
        "parity": parity,
    }


def three_stage_leg(args, c2):
    """The same 2 GiB workload through the three_stage pipeline (bounded
    labeling, part_private masses, cluster_merge exact sums, divide)."""
    run = c2.measure("three_stage", "cluster_merge", args.steps, args.warmup, "unroll_vector_bounded")
    n, d = c2.n, c2.d
    by = run["by_name"]
    hot, hot_ms, _ = hot_kernel(by, ["LabelingUnrollVectorBounded", "LabelingUnrollVector"])
    bytes_iter = n * (8 * d + 12)  # labeling 4d+4, masses 4, centroid update 4d+4
    ach_iter = bytes_iter / (run["ms_per_step"] * 1e-3) / 1e9
    out = {"pipeline": "three_stage", "labeling": "unroll_vector_bounded", "mass_update": "part_private",
           "centroid_update": "cluster_merge", "value": 1000.0 / run["ms_per_step"], "unit": "iter/s",
           "ms_per_step": run["ms_per_step"],
           "kernels_ms": {kk: round(float(np.mean(v)), 4) for kk, v in by.items()},
           "roofline": {"bound": "hbm", "achieved": ach_iter, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach_iter / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "whole iteration (labeling + mass + centroid update + divide)",
                        "bytes_per_iteration": bytes_iter, "labeling_kernel": hot, "labeling_ms": hot_ms}}
    if not args.no_cpu:
        from oracle import oracle as O
        out["parity"] = c2.full_parity(run, O)
        t = time.perf_counter()
        O.lloyd3_mt(c2.pts, c2.init, 1, c2.threads)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "iter/s", "cores": c2.threads, "kind": "port",
                               "sample": "1 full iteration of the same workload, three-stage restatement "
                                         "(oracle/kmeans_oracle.c ora_lloyd3_mt_f32: label, mass and centroid "
                                         "passes), %d threads" % c2.threads, **host_info()}
    return out


def buffered_leg(args, c2):
    """single_stage_buffered through the C++ strategy (clk_kmeans_run): the
    2 GiB streamed from pinned host memory in 256 MiB chunks each iteration,
    labels read back chunkwise; against the measured pinned host->device copy
    rate (its roofline)."""
    import clkmeans
    from clkmeans import hip
    chunk = 256 << 20
    # the host link: pinned hipMemcpyAsync host->device of 256 MiB chunks
    src = hip.PinnedArray((chunk // 4,), np.float32)
    dst = hip.DeviceArray((chunk // 4,), np.float32)
    s = hip.Stream()
    hip.check(hip.lib().hipMemcpyAsync(dst.ptr, src.ptr, chunk, hip.hipMemcpyHostToDevice, s.handle), "H2D")
    s.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for _ in range(8):
        hip.check(hip.lib().hipMemcpyAsync(dst.ptr, src.ptr, chunk, hip.hipMemcpyHostToDevice, s.handle), "H2D")
    e1.record(s)
    e1.synchronize()
    link = 8 * chunk / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    del src, dst
    iters = args.buffered_iters
    conf = ("[kmeans]\nclusters = %d\npipeline = single_stage_buffered\niterations = %d\n"
            "[kmeans.fused]\nstrategy = cluster_merge_bounded\n" % (c2.k, iters))
    c, lab, m, it, ns = clkmeans.kmeans_run(conf, c2.pts, c2.init, buffer_size=chunk)
    ms = ns * 1e-6 / max(it, 1)
    link_bytes = c2.n * 4 * c2.d  # points in (the labels go out on the other direction of the link)
    ach = link_bytes / (ms * 1e-3) / 1e9
    out = {"pipeline": "single_stage_buffered", "strategy": "cluster_merge_bounded", "chunk_bytes": chunk,
           "iterations": it, "value": 1000.0 / ms, "unit": "iter/s", "ms_per_step": ms,
           "roofline": {"bound": "host link (pinned H2D)", "achieved": ach, "peak": link, "unit": "GB/s",
                        "frac": ach / link, "traffic": None, "bytes_per_iteration": link_bytes,
                        "d2h_label_bytes_per_iteration": c2.n * 4,
                        "peak_kind": "pinned hipMemcpyAsync host->device, 256 MiB chunks, measured in this run"}}
    if not args.no_cpu:
        from oracle import oracle as O
        rc, rl, rm = O.lloyd_mt(c2.pts, c2.init, it, c2.threads)
        out["parity"] = {"labels_equal_oracle": bool(np.array_equal(lab, rl)),
                         "masses_equal_oracle": bool(np.array_equal(m, rm)),
                         "centroid_max_rel_diff_vs_fp64_loop": float(np.nanmax(np.abs(c - rc) /
                                                                               np.maximum(np.abs(rc), 1e-30)))}
        out["cpu_baseline"] = {"note": "the CPU has no host-link stage: the headline's cpu_baseline "
                                       "(the same workload in host memory) applies"}
    return out


# --------------------------------------------------------------------------
# C1: 2^20 points, 30 iterations in full


def c1_leg(args):
    import clkmeans
    from oracle import oracle as O
    n, d, k, iters = 1 << 20, 4, 10, 30
    threads = cpu_threads()
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=50.0, seed=0, threads=threads)
    init = np.ascontiguousarray(pts[:, :k])
    out = {"workload": "N=2^20 d=4 k=10, sigma=1 blobs in [-50,50]^4, seed 0, first-x init, 30 iterations",
           "note": "20 MiB per iteration sits in the 256 MB Infinity Cache: the HBM fraction is reported "
                   "for reference only"}
    gc, gl, gm = O.lloyd(pts, init, iters, mode=O.MODE_GRID)
    for pipeline, conf_extra in (("single_stage", "[kmeans.fused]\nstrategy = cluster_merge_bounded\n"),
                                 ("three_stage", "[kmeans.labeling]\nstrategy = unroll_vector_bounded\n")):
        conf = "[kmeans]\nclusters = %d\npipeline = %s\niterations = %d\n%s" % (k, pipeline, iters, conf_extra)
        clkmeans.kmeans_run(conf, pts, init)  # warm-up (first launches)
        c, lab, m, it, ns = clkmeans.kmeans_run(conf, pts, init)
        ms = ns * 1e-6 / it
        bytes_iter = n * (4 * d + 4) if pipeline == "single_stage" else n * (8 * d + 12)
        ach = bytes_iter / (ms * 1e-3) / 1e9
        t = time.perf_counter()
        if pipeline == "single_stage":
            O.lloyd_mt(pts, init, iters, threads)
        else:
            O.lloyd3_mt(pts, init, iters, threads)
        dt = time.perf_counter() - t
        out[pipeline] = {
            "value": 1000.0 / ms, "unit": "iter/s", "ms_per_step": ms, "iterations": it,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": None, "bytes_per_iteration": bytes_iter,
                         "kernel": "whole iteration (the strategy's TotalTime)"},
            "cpu_baseline": {"value": iters / dt, "unit": "iter/s", "cores": threads, "kind": "port",
                             "sample": "all 30 iterations, %s restatement (%s), %d threads" %
                                       (pipeline, "ora_lloyd_mt_f32" if pipeline == "single_stage"
                                        else "ora_lloyd3_mt_f32", threads), **host_info()},
            "parity": {"labels_equal_oracle_grid_loop": bool(np.array_equal(lab, gl)),
                       "masses_equal": bool(np.array_equal(m, gm)),
                       "centroids_bit_equal_oracle_grid_loop": bool(np.array_equal(c, gc))},
        }
    return out


# --------------------------------------------------------------------------
# C3: dense distances on MFMA


def dense(args):
    """The dense-distance configuration (d=64, k=1024) on the MFMA labeling."""
    import clkmeans
    from clkmeans import hip

    log2n = 24
    n, d, k = 1 << log2n, 64, 1024
    steps = args.dense_steps
    threads = cpu_threads()
    t0 = time.time()
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=0.5, box=1.0, seed=0, threads=threads)
    log("generated %d x %d points in %.1fs" % (n, d, time.time() - t0))
    init = np.ascontiguousarray(pts[:, :k])
    ctx = clkmeans.Context(0)
    stream = hip.Stream()
    s = stream.handle
    dp = hip.DeviceArray.from_host(pts)
    cen = [hip.DeviceArray.from_host(init), hip.DeviceArray((d, k), np.float32)]
    sums = hip.DeviceArray((d, k), np.float64)
    masses = hip.DeviceArray((k,), np.uint32)
    labels = hip.DeviceArray((n,), np.uint32)
    lcfg = ctx.config(clkmeans.STAGE_LABELING, args.labeling)
    mcfg = ctx.config(clkmeans.STAGE_MASS_UPDATE, "part_local")
    ccfg = ctx.config(clkmeans.STAGE_CENTROID_UPDATE, "cluster_merge")
    ctx.sum_grid_from_points(d, n, dp, stream=s)  # pinned once (setup)
    pinned = args.labeling.startswith("mfma") and not args.no_pin
    if pinned:  # the 16-bit operand copy of the points, once (setup)
        ctx.labeling_pin_points(args.labeling, d, n, dp, stream=s)
    stream.synchronize()

    def iteration(lab_cfg):
        old, new = cen[0], cen[1]
        ctx.fill_zero(masses, stream=s)
        ctx.fill_zero(sums, stream=s)
        ctx.labeling(lab_cfg, d, n, k, dp, old, labels, stream=s)
        ctx.mass_update(mcfg, n, k, labels, masses, stream=s)
        ctx.centroid_update(ccfg, d, n, k, dp, sums, labels, masses, stream=s)
        ctx.divide(d, k, sums, masses, new, stream=s)
        cen[0], cen[1] = new, old

    for _ in range(args.warmup):
        iteration(lcfg)
    stream.synchronize()
    # the step time (events around the K iterations only), then K more
    # iterations with every kernel bracketed by events for the roofline
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for it in range(steps):
        ctx.set_iteration(it)
        iteration(lcfg)
    e1.record(stream)
    e1.synchronize()
    ms_per_step = e0.elapsed_ms(e1) / steps
    ctx.clear_records()
    ctx.profiling(True)
    p0, p1 = hip.Event(), hip.Event()
    p0.record(stream)
    for it in range(steps):
        ctx.set_iteration(steps + it)
        iteration(lcfg)
    p1.record(stream)
    p1.synchronize()
    ctx.profiling(False)
    ms_instrumented = p0.elapsed_ms(p1) / steps
    by = kernel_times(ctx)
    log("dense kernels (avg ms): " + ", ".join("%s=%.4f x%d" % (kk, np.mean(v), len(v)) for kk, v in by.items()))
    hot, hot_ms, _ = hot_kernel(by, ["LabelingMfmaF16" if args.labeling == "mfma_fp16" else "LabelingMfmaBF16",
                                     "LabelingGlobal", "LabelingChunked"])
    flops = 2.0 * n * k * d
    achieved = flops / (hot_ms * 1e-3) / 1e12
    refined = ctx.labeling_refined_points() if args.labeling.startswith("mfma") else None

    # one more iteration, untimed, for the checks
    c_in = cen[0].to_host()
    iteration(lcfg)
    stream.synchronize()
    lab, m, c_out = labels.to_host(), masses.to_host(), cen[0].to_host()
    # the plain FP32 labeling (unroll_vector) of the same centroids: same labels
    pcfg = ctx.config(clkmeans.STAGE_LABELING, "unroll_vector")
    c_dev = hip.DeviceArray.from_host(c_in)
    ctx.clear_records()
    ctx.profiling(True)
    ctx.labeling(pcfg, d, n, k, dp, c_dev, labels, stream=s)
    stream.synchronize()
    ctx.profiling(False)
    plain_ms = float(sum(ms for _, _, ms in ctx.records()))
    plain_equal = bool(np.array_equal(labels.to_host(), lab))
    ctx.sum_grid_clear()
    ctx.labeling_unpin_points()
    ctx.close()
    del dp, cen, sums, masses, labels, c_dev
    out = {
        "workload": "dense Lloyd, N=2^%d d=64 k=1024 float32 feature-major, sigma=0.5 blobs in [-1,1]^64, seed 0, "
                    "first-x init; three_stage: %s labeling%s, part_local masses, feature-slice exact sums, "
                    "FP64 divide" % (log2n, args.labeling,
                                     " (points pinned once as its 16-bit operand copy)" if pinned else ""),
        "value": 1000.0 / ms_per_step, "unit": "iter/s", "ms_per_step": ms_per_step, "steps": steps,
        "dtype": "f16-mfma+f32-exact" if args.labeling == "mfma_fp16" else
                 "bf16-mfma+f32-exact" if args.labeling == "mfma_bf16" else "f32",
        "kernels_ms": {kk: round(float(np.mean(v)), 4) for kk, v in by.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA16_PEAK_TFLOPS, "traffic": None, "kernel": hot,
                     "kernel_ms": hot_ms, "flops_per_launch": flops,
                     "kernel_ms_source": "the %d iterations after the timed ones, every kernel bracketed by HIP "
                                         "events (%.4f ms per step with those events)" % (steps, ms_instrumented)},
        "refined_points_last_call": refined,
        "plain_fp32_labeling_ms": plain_ms,
        "parity": {"plain_fp32_labels_equal": plain_equal},
    }
    if not args.no_cpu:
        # The CPU baseline leg, and the oracle as the checker: a label sample
        # of the last labeling; the centroids it produced against the
        # oracle's exact grid sums (bit-exact) and against FP64 sums.
        from oracle import oracle as O
        idx = np.sort(np.random.default_rng(0).choice(n, size=4096, replace=False))
        out["parity"]["labels_sample"] = int(idx.size)
        out["parity"]["labels_equal"] = bool(np.array_equal(lab[idx],
                                                            O.label(np.ascontiguousarray(pts[:, idx]), c_in)))
        ref = O.sums_grid(pts, lab, k)
        with np.errstate(invalid="ignore", divide="ignore"):
            out["parity"]["centroids_equal_oracle_grid_sums"] = bool(np.array_equal(c_out,
                                                                                    O.divide_from_f64(ref, m)))
            ref64 = O.sums_f64(pts, lab, k) / m.astype(np.float64)
        out["parity"]["centroid_max_rel_err_vs_fp64"] = centroid_err(c_out, ref64, pts, lab)
        ns = 1 << 16
        sub = np.ascontiguousarray(pts[:, :ns])
        t = time.perf_counter()
        O.lloyd_mt(sub, init, 1, threads)
        dt = (time.perf_counter() - t) * (n / ns)
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "iter/s", "cores": threads, "kind": "port",
                               "sample": "one fused cluster_merge iteration (ora_lloyd_mt_f32) over the first 2^16 "
                                         "of the 2^%d points against the same 1024 centroids, %d threads, scaled by "
                                         "N/2^16" % (log2n, threads), **host_info()}
    return out


# --------------------------------------------------------------------------
# C4: the batch engine


def batch(args):
    """Many uneven jobs: the batch engine against the serial per-job loop."""
    import clkmeans
    from clkmeans import hip

    iters = 25
    mix = clkmeans.job_mix(args.jobs, seed=0)
    threads = cpu_threads()
    ctx = clkmeans.Context(0)
    stream = hip.Stream()
    s = stream.handle
    probe = clkmeans.Batch(ctx, 1, 1 << 20, 32, 256)
    need = sum(probe.job_bytes(n, d, k) for n, d, k, _ in mix)
    probe.close()
    arena = int(need * 1.02) + (64 << 20)
    # Tuning (outside the timed run): the autotuner on a copy of the jobs
    # after their second iteration, when the keep test that later iterations
    # run is on (the first iteration's full scans favour smaller tiles); the
    # tiling changes no result.  The timed engine then starts every job from
    # init.
    tuned = args.tile_work
    tune_note = "given"
    if not tuned:
        t0 = time.time()
        sub = mix
        tb = clkmeans.Batch(ctx, 1, 1 << 20, 32, 256)
        need_t = sum(tb.job_bytes(n, d, k) for n, d, k, _ in sub)
        tb.close()
        tun = clkmeans.Batch(ctx, len(sub), int(need_t * 1.02) + (64 << 20), 32, 256, labeling=args.batch_labeling)
        for n, d, k, seed in sub:
            pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed, threads=threads)
            tun.add(pts, np.ascontiguousarray(pts[:, :k]), iters, stream=s)
        for _ in range(2):
            tun.step(s)
        tuned = tun.autotune(s)
        tun.close()
        tune_note = "autotuned on a copy of the jobs after 2 iterations"
        log("autotuned tile_work = %d (%d jobs, %.1fs)" % (tuned, len(sub), time.time() - t0))
    eng = clkmeans.Batch(ctx, len(mix), arena, 32, 256, tile_work=tuned, labeling=args.batch_labeling)
    ser = clkmeans.Batch(ctx, len(mix), arena, 32, 256)
    t0 = time.time()
    slots = []
    checks = {}
    small = [j for j, (n, d, k, _) in enumerate(mix) if n * k * d <= 2e7]
    check_ids = set(np.random.default_rng(1).choice(small, size=min(12, len(small)), replace=False).tolist())
    cpu_sub = {}
    for j, (n, d, k, seed) in enumerate(mix):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed, threads=threads)
        init = np.ascontiguousarray(pts[:, :k])
        slots.append((eng.add(pts, init, iters, stream=s), ser.add(pts, init, iters, stream=s)))
        if j in check_ids:
            checks[j] = (pts, init)
        if j < 64:
            cpu_sub[j] = pts
    log("generated and loaded %d jobs (%.2f GB of points) in %.1fs" %
        (len(mix), sum(n * d * 4 for n, d, _, _ in mix) / 1e9, time.time() - t0))
    points = sum(n for n, _, _, _ in mix)
    bytes_per_iter = sum(n * (4 * d + 4) for n, d, _, _ in mix)
    dist_flops = sum(2.0 * n * k * d for n, d, k, _ in mix)
    # the engine labels jobs with k >= 16 and k*d >= 3072 on the matrix cores
    # (MFMA labelings), the others with the FP32 scan
    on_mfma = lambda d, k: args.batch_labeling.startswith("mfma") and k >= 16 and k * d >= 3072
    mfma_flops = sum(2.0 * n * k * d for n, d, k, _ in mix if on_mfma(d, k))
    fp32_flops = dist_flops - mfma_flops

    # the batch engine: graph capture on the first step, replays timed
    e0, e1 = hip.Event(), hip.Event()
    stream.synchronize()
    ctx.clear_records()
    ctx.profiling(True)
    e0.record(stream)
    for _ in range(iters):
        eng.step(s)
    e1.record(stream)
    e1.synchronize()
    ctx.profiling(False)
    batch_ms = e0.elapsed_ms(e1)
    step_ms = [ms for name, _, ms in ctx.records() if name == "BatchStep"]
    if step_ms:
        log("batch step ms: " + " ".join("%.2f" % x for x in step_ms))
    states, active = eng.poll(s)
    # the plain per-job loop on the same GPU
    stream.synchronize()
    w0 = time.perf_counter()
    if not args.no_serial:
        ser.run_serial(s)
    stream.synchronize()
    serial_ms = (time.perf_counter() - w0) * 1e3 if not args.no_serial else None

    # parity: every job's bits equal the per-job loop's (exact sums); a
    # sample against the oracle
    lab_equal, mass_equal, cen_equal, iters_equal = 0, 0, 0, 0
    sample_out = {}
    for j, (a, b_) in enumerate(slots):
        ca, la, ma, ita, _ = eng.take(a, s)
        if j in checks:
            sample_out[j] = (ca, la, ma)
        if args.no_serial:
            continue
        cb, lb, mb, itb, _ = ser.take(b_, s)
        lab_equal += int(np.array_equal(la, lb))
        mass_equal += int(np.array_equal(ma, mb))
        cen_equal += int(np.array_equal(ca.view(np.uint32), cb.view(np.uint32)))
        iters_equal += int(ita == itb == iters)
    eng.close()
    ser.close()
    ctx.close()

    it_ms = float(np.mean(step_ms)) if step_ms else batch_ms / iters
    ach_hbm = bytes_per_iter / (it_ms * 1e-3) / 1e9
    ach_flops = dist_flops / (it_ms * 1e-3) / 1e12
    hbm_floor_ms = bytes_per_iter / (HBM_PEAK_GBS * 1e9) * 1e3
    fp32_floor_ms = fp32_flops / (FP32_PEAK_TFLOPS * 1e12) * 1e3
    mfma_floor_ms = mfma_flops / (MFMA16_PEAK_TFLOPS * 1e12) * 1e3
    compute_floor_ms = fp32_floor_ms + mfma_floor_ms
    bound = "hbm" if hbm_floor_ms >= compute_floor_ms else "compute (fp32 vector + mfma)"
    floor_ms = max(hbm_floor_ms, compute_floor_ms)
    out = {
        "workload": "%d jobs: N log-uniform [2^10,2^18], d in {3,8,16,32}, k log-uniform [4,256] capped at N/16, "
                    "sigma=1 blobs in [-10,10]^d, per-job seeds, first-x init, %d iterations; batch engine (size-"
                    "class persistent launches per iteration, HIP graph), %s labeling"
                    % (len(mix), iters, args.batch_labeling),
        "value": len(mix) / (batch_ms * 1e-3), "unit": "solves/s", "ms_per_step": batch_ms / iters,
        "jobs": len(mix), "points": points, "labeling": args.batch_labeling, "tile_work": tuned,
        "tile_work_choice": tune_note,
        "dtype": "f32" if args.batch_labeling == "unroll_vector" else
                 args.batch_labeling.split("_")[1] + "-mfma+f32-exact",
        "serial_per_job_loop": {"solves_per_s": len(mix) / (serial_ms * 1e-3), "total_ms": serial_ms}
                               if serial_ms else None,
        "speedup_vs_serial": serial_ms / batch_ms if serial_ms else None,
        "roofline": {"bound": bound, "achieved": ach_hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": floor_ms / it_ms, "traffic": None,
                     "kernel": "one batch iteration (size-class tile kernels, merge, finish)",
                     "bytes_per_iteration": bytes_per_iter, "iteration_ms": it_ms,
                     "floor_ms": {"hbm": hbm_floor_ms, "fp32_classes_on_vector_peak": fp32_floor_ms,
                                  "mfma_classes_on_matrix_peak": mfma_floor_ms},
                     "flops_per_iteration": {"fp32_classes": fp32_flops, "mfma_classes": mfma_flops},
                     "achieved_TFLOPs": ach_flops,
                     "note": "floors: points + labels bytes at 8 TB/s; 2*n*k*d distance flops of the FP32-scan "
                             "jobs at the 157.3 TFLOP/s vector peak plus those of the MFMA jobs at 2.5 PFLOP/s "
                             "(compute and memory overlap: the larger floor is the bound); frac = floor / "
                             "iteration time. The per-point labeling, run sort and LDS accumulation issue "
                             "rates are the limit (PMC in profiles/batch_step_c4mix_4096jobs.md)"},
        "parity": {"jobs_labels_equal_serial": lab_equal, "jobs_masses_equal_serial": mass_equal,
                   "jobs_centroids_bit_equal_serial": cen_equal, "jobs_iterations_equal_serial": iters_equal,
                   "jobs_done": int((states == clkmeans.JOB_DONE).sum()), "still_active": int(active)},
    }
    if not args.no_cpu:
        # The CPU baseline leg (data generated beforehand, outside the timed
        # region) and the oracle as the checker of the sampled jobs.
        from oracle import oracle as O
        ok, exact, fp64_err = True, True, 0.0
        for j, (ca, la, ma) in sample_out.items():
            pts, init = checks[j]
            rc, rl, rm = O.lloyd(pts, init, iters, mode=O.MODE_GRID)
            ok &= bool(np.array_equal(la, rl)) and bool(np.array_equal(ma, rm))
            exact &= bool(np.array_equal(ca, rc))
            with np.errstate(invalid="ignore", divide="ignore"):
                ref64 = O.sums_f64(pts, rl, init.shape[1]) / rm.astype(np.float64)
            fp64_err = max(fp64_err, centroid_err(ca, ref64, pts, rl))
        out["parity"].update({"oracle_sample_jobs": len(checks), "oracle_sample_labels_masses_equal": ok,
                              "oracle_sample_centroids_bit_equal_grid_loop": exact,
                              "oracle_sample_centroid_max_rel_err_vs_fp64": fp64_err})
        sub = mix[:64]
        t = time.perf_counter()
        for j, (n, d, k, seed) in enumerate(sub):
            p = cpu_sub[j]
            O.lloyd_mt(p, np.ascontiguousarray(p[:, :k]), iters, threads)
        dt = time.perf_counter() - t
        w_sub = sum(n * k * d for n, d, k, _ in sub)
        w_all = sum(n * k * d for n, d, k, _ in mix)
        out["cpu_baseline"] = {"value": len(mix) / (dt * w_all / w_sub), "unit": "solves/s", "cores": threads,
                               "kind": "port",
                               "sample": "the first 64 jobs, 25 iterations each, fused cluster_merge restatement "
                                         "(ora_lloyd_mt_f32, %d threads; data generated before timing), scaled by "
                                         "total n*k*d work" % threads, **host_info()}
    return out


# --------------------------------------------------------------------------
# C5: serving


def serve(args):
    """Poisson arrivals of the batched mix plus 2% large jobs, continuous
    batching (clk_serve_run) against the serial plain-FP32 per-job loop
    (clk_serve_serial) on the same arrivals; with the points already in HBM
    (value), copied in from pinned host memory on admission, and a
    latency-versus-rate sweep."""
    import clkmeans
    from clkmeans import hip

    threads = cpu_threads()
    count = args.serve_jobs
    rng = np.random.default_rng(7)
    mix = clkmeans.job_mix(count, seed=1)
    large = set(rng.choice(count, size=int(round(args.large_frac * count)), replace=False).tolist())
    ctx = clkmeans.Context(0)
    stream = hip.Stream()
    t0 = time.time()
    host, dev, inits, specs = [], [], [], []
    large_data = []  # a few distinct large data sets, shared by the large jobs (read only)
    for j, (n, d, k, seed) in enumerate(mix):
        if j in large:
            n, d, k = 1 << 24, 32, 256
            if len(large_data) < 4:
                pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=900 + len(large_data),
                                                 threads=threads)
                p = hip.PinnedArray(pts.shape, np.float32)
                p.array[...] = pts
                large_data.append((p, hip.DeviceArray.from_host(pts), np.ascontiguousarray(pts[:, :k])))
                del pts
            p, dp, init = large_data[j % len(large_data)]
        else:
            raw, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed, threads=threads)
            p = hip.PinnedArray(raw.shape, np.float32)
            p.array[...] = raw
            dp, init = hip.DeviceArray.from_host(raw), np.ascontiguousarray(raw[:, :k])
        host.append(p)
        dev.append(dp)
        inits.append(init)
        specs.append((n, d, k))
    log("generated %d jobs (%d large, %.2f GB of points) in %.1fs" %
        (count, len(large), sum(n * d * 4 for n, d, _ in specs) / 1e9, time.time() - t0))
    probe = clkmeans.Batch(ctx, 1, 1 << 20, 32, 256)
    big_bytes = probe.job_bytes(1 << 24, 32, 256) if large else probe.job_bytes(1 << 18, 32, 256)
    probe.close()

    def requests(ids, arrivals, resident):
        return [clkmeans.ServeRequest(dev[j] if resident else host[j].array, inits[j], t,
                                      max_iterations=50, tolerance=1e-4, want_labels=False)
                for j, t in zip(ids, arrivals)]

    def run(ids, rate, resident, seed=3):
        arrivals = clkmeans.poisson_arrivals(len(ids), rate, seed=seed)
        reqs = requests(ids, arrivals, resident)
        eng = clkmeans.Batch(ctx, 1024, int(big_bytes * 12 + (8 << 30)), 32, 256, tile_work=args.tile_work,
                             labeling=args.batch_labeling)
        st = clkmeans.serve(eng, reqs, stream.handle)
        eng.close()
        sreqs = requests(ids, arrivals, resident)
        ss = clkmeans.serve_serial(ctx, sreqs, stream.handle)
        same_iters = sum(int(a.iterations == b.iterations) for a, b in zip(reqs, sreqs))
        same_mass = sum(int(np.array_equal(a.out_masses, b.out_masses)) for a, b in zip(reqs, sreqs))
        same_cen = sum(int(np.array_equal(a.out_centroids.view(np.uint32), b.out_centroids.view(np.uint32)))
                       for a, b in zip(reqs, sreqs))
        lat_large = [r.latency for r, j in zip(reqs, ids) if j in large] or [0.0]
        return {
            "jobs": len(ids), "rate": rate,
            "served": {"solves_per_s": st["solves_per_s"], "makespan_s": st["makespan"], "p50_s": st["p50"],
                       "p99_s": st["p99"], "mean_latency_s": st["mean_latency"],
                       "iterations_launched": int(st["steps"]), "max_active": int(st["max_active"]),
                       "large_job_latency_s": {"min": min(lat_large), "max": max(lat_large)}},
            "serial_per_job_loop": {"solves_per_s": ss["solves_per_s"], "makespan_s": ss["makespan"],
                                    "p50_s": ss["p50"], "p99_s": ss["p99"]},
            "speedup_vs_serial": st["solves_per_s"] / ss["solves_per_s"],
            "parity": {"jobs": len(ids), "jobs_same_iterations_as_serial": same_iters,
                       "jobs_same_masses_as_serial": same_mass, "jobs_same_centroid_bits_as_serial": same_cen},
        }

    ids = list(range(count))
    res = run(ids, args.rate, True)
    log("serve: resident done")
    pcie = run(ids, args.rate, False)
    log("serve: pcie-inclusive done")
    sweep = []
    if args.sweep_jobs:
        sub = ids[:args.sweep_jobs]
        for rate in (300.0, 1000.0, 3000.0, 10000.0):
            r = run(sub, rate, True, seed=11)
            sweep.append({"rate": rate, "jobs": len(sub),
                          "served": {kk: r["served"][kk] for kk in ("solves_per_s", "p50_s", "p99_s")},
                          "serial": {kk: r["serial_per_job_loop"][kk] for kk in ("solves_per_s", "p50_s", "p99_s")},
                          "jobs_same_centroid_bits_as_serial": r["parity"]["jobs_same_centroid_bits_as_serial"]})
            log("serve: rate %.0f done" % rate)
    ctx.close()
    sr = res["served"]
    return {
        "workload": "%d jobs, Poisson arrivals at %.0f jobs/s: the batched mix (N log-uniform [2^10,2^18], d in "
                    "{3,8,16,32}, k log-uniform [4,256] capped at N/16) with %d large jobs (N=2^24, d=32, k=256; "
                    "4 distinct data sets); each until max shift < 1e-4 or 50 iterations; points already in HBM"
                    % (count, args.rate, len(large)),
        "value": sr["solves_per_s"], "unit": "solves/s", "labeling": args.batch_labeling,
        "dtype": "f32" if args.batch_labeling == "unroll_vector" else
                 args.batch_labeling.split("_")[1] + "-mfma+f32-exact",
        **res,
        "pcie_inclusive": pcie,
        "latency_vs_rate": sweep,
        "cpu_baseline": {"note": "per-job CPU cost as in the batch leg's cpu_baseline (the same job mix)"},
    }


# --------------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--d", type=int, default=4)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--strategy", default="cluster_merge_bounded",
                    help="fused strategy of the headline; other than cluster_merge, the exhaustive cluster_merge "
                         "pass is timed in the same run too (--no-compare skips it)")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="the headline only (no sub-workloads)")
    ap.add_argument("--workload", default=None,
                    choices=["fused", "three_stage", "buffered", "c1", "dense", "batch", "serve"],
                    help="run one workload alone and print its object")
    ap.add_argument("--buffered-iters", type=int, default=5)
    ap.add_argument("--dense-steps", type=int, default=10)
    ap.add_argument("--serve-jobs", type=int, default=1000, help="serve workload: jobs in the arrival stream")
    ap.add_argument("--sweep-jobs", type=int, default=200,
                    help="serve workload: jobs per rate of the latency sweep (0: no sweep)")
    ap.add_argument("--rate", type=float, default=10000.0,
                    help="serve workload: Poisson arrival rate (jobs/s; the default saturates both loops)")
    ap.add_argument("--large-frac", type=float, default=0.02,
                    help="serve workload: share of large jobs (N=2^24, d=32, k=256)")
    ap.add_argument("--jobs", type=int, default=4096, help="batch workload: number of jobs")
    ap.add_argument("--no-serial", action="store_true",
                    help="batch workload: skip the plain per-job loop (for profiling the batch kernels)")
    ap.add_argument("--tile-work", type=int, default=0,
                    help="batch and serve workloads: points per cluster in a tile (0: batch autotuned, serve "
                         "the engine's default)")
    ap.add_argument("--batch-labeling", default="mfma_bf16", help="batch and serve: labeling strategy")
    ap.add_argument("--labeling", default="mfma_fp16", help="dense workload: labeling strategy")
    ap.add_argument("--no-pin", action="store_true",
                    help="dense workload: label from the FP32 points instead of the pinned 16-bit operand copy")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scales rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import clkmeans  # noqa: F401
    from clkmeans import hip

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a, b = hip.DeviceArray((nbytes,), np.uint8), hip.DeviceArray((nbytes,), np.uint8)
        e0, e1 = hip.Event(), hip.Event()
        hip.check(hip.lib().hipMemcpy(b.ptr, a.ptr, nbytes, hip.hipMemcpyDeviceToDevice), "hipMemcpy D2D")
        e0.record()
        for _ in range(3):
            hip.check(hip.lib().hipMemcpyAsync(b.ptr, a.ptr, nbytes, hip.hipMemcpyDeviceToDevice, None),
                      "hipMemcpyAsync D2D")
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / 3
        print(json.dumps({"calibration": "hipMemcpy D2D", "bytes_each": nbytes, "ms_each": ms,
                          "read_plus_write_GBps": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return

    alone = {"c1": c1_leg, "dense": dense, "batch": batch, "serve": serve}
    if args.workload in alone:
        print(json.dumps(alone[args.workload](args)), flush=True)
        return

    w0 = time.time()
    c2 = C2(args)
    if args.workload == "three_stage":
        out = three_stage_leg(args, c2)
    elif args.workload == "buffered":
        out = buffered_leg(args, c2)
    else:
        out = fused_headline(args, c2)
        log("headline done (%.0fs)" % (time.time() - w0))
        if not args.no_extra and args.workload is None:
            out["three_stage"] = three_stage_leg(args, c2)
            log("three_stage done (%.0fs)" % (time.time() - w0))
            out["buffered"] = buffered_leg(args, c2)
            log("buffered done (%.0fs)" % (time.time() - w0))
    c2.close()
    del c2
    if args.workload is None and not args.no_extra:
        for name, fn in (("c1", c1_leg), ("dense", dense), ("batch", batch), ("serve", serve)):
            out[name] = fn(args)
            log("%s done (%.0fs)" % (name, time.time() - w0))
    out["bench_wall_s"] = time.time() - w0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
