"""Lloyd-iteration throughput on one MI355X.

Workload (default): N=2^27 points, d=4 float32 features (2 GiB, feature-major),
k=10, Gaussian blobs sigma=1 with centres uniform in [-50,50]^4, seed 0,
init = first 10 points; one step = one Lloyd iteration of the single_stage
pipeline (zero accumulators, fused label + mass + feature-sum pass, FP64
merge, FP64 divide, swap).  Points are resident in HBM before timing starts.

Prints one JSON line: iter/s, the fused kernel's achieved HBM bandwidth
against the 8 TB/s peak (algorithmic bytes N*(4d+4) per launch over its
HIP-event-timed average duration), and the multi-threaded CPU restatement
timed on the host cores.

--workload batch: 4096 independent jobs, N log-uniform in [2^10, 2^18], d
uniform in {3, 8, 16, 32}, k log-uniform in [4, 256] capped at N/16, blobs
sigma=1 with centres uniform in [-10,10]^d (seed per job), first-x init, 25
iterations each; value = solves/s of the batch engine (one persistent launch
per iteration over all jobs, replayed from a HIP graph) next to the serial
per-job loop (clk_lloyd_step per job and iteration) on the same GPU.

--workload dense: N=2^24 points, d=64, k=1024, blobs sigma=0.5 with centres
uniform in [-1,1]^64, init = first 1024 points; one step = one three_stage
iteration with MFMA labeling (mfma_fp16 by default, --labeling), masses,
feature-slice FP64 centroid sums and the FP64 divide.  The roofline is the
labeling kernel's x·c^T rate (2*N*k*d flops per launch) against the 2.5
PFLOP/s dense FP16/BF16 MFMA peak.
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
METRIC = "Lloyd iter/s+%HBM (4f/10c 2GiB); served/batched vs plain solves/s @tol; p99"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--d", type=int, default=4)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--pipeline", default="single_stage", choices=["single_stage", "three_stage"])
    ap.add_argument("--strategy", default="cluster_merge_bounded",
                    help="fused strategy; other than cluster_merge, the exhaustive cluster_merge pass is timed "
                         "in the same run too (--no-compare skips it)")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--stage-labeling", default="unroll_vector", help="three_stage pipeline: labeling strategy")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="fused", choices=["fused", "dense", "batch", "serve"])
    ap.add_argument("--serve-jobs", type=int, default=1000, help="serve workload: jobs in the arrival stream")
    ap.add_argument("--rate", type=float, default=10000.0,
                    help="serve workload: Poisson arrival rate (jobs/s; the default saturates both loops)")
    ap.add_argument("--large-frac", type=float, default=0.02,
                    help="serve workload: share of large jobs (N=2^24, d=32, k=256)")
    ap.add_argument("--no-serial", action="store_true",
                    help="batch workload: skip the serial per-job loop (for profiling the batch kernels)")
    ap.add_argument("--jobs", type=int, default=4096, help="batch workload: number of jobs")
    ap.add_argument("--tile-work", type=int, default=0,
                    help="serve workload: points per cluster in a tile (0: the engine's default)")
    ap.add_argument("--batch-labeling", default="mfma_bf16",
                    help="batch workload: labeling strategy of the batch engine")
    ap.add_argument("--labeling", default="mfma_fp16", help="dense workload: labeling strategy")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scales rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import clkmeans
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

    if args.workload == "dense":
        return dense(args)
    if args.workload == "serve":
        return serve(args)
    if args.workload == "batch":
        return batch(args)

    n, d, k = 1 << args.log2n, args.d, args.k
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=50.0, seed=0, threads=threads)
    log("generated %d x %d points in %.1fs" % (n, d, time.time() - t0))
    init = np.ascontiguousarray(pts[:, :k])  # first_x

    ctx = clkmeans.Context(0)
    stream = hip.Stream()
    dp = hip.DeviceArray.from_host(pts)
    cen = [hip.DeviceArray.from_host(init), hip.DeviceArray((d, k), np.float32)]
    sums = hip.DeviceArray((d, k), np.float64)
    masses = hip.DeviceArray((k,), np.uint32)
    labels = hip.DeviceArray((n,), np.uint32)
    lcfg = ctx.config(clkmeans.STAGE_LABELING, args.stage_labeling)
    mcfg = ctx.config(clkmeans.STAGE_MASS_UPDATE, "part_private")
    ccfg = ctx.config(clkmeans.STAGE_CENTROID_UPDATE, "cluster_merge")
    s = stream.handle
    # The sum grid of the data set, pinned once (setup, like the upload; the
    # strategies do the same): every iteration's sums are exact on it.
    ctx.sum_grid_from_points(d, n, dp, stream=s)
    stream.synchronize()

    def iteration(fcfg):
        old, new = cen[0], cen[1]
        if args.pipeline == "single_stage":
            # zero + fused pass + FP64 divide: one launch
            ctx.lloyd_step(fcfg, d, n, k, dp, old, new, sums, labels, masses, stream=s)
        else:
            ctx.fill_zero(masses, stream=s)
            ctx.fill_zero(sums, stream=s)
            ctx.labeling(lcfg, d, n, k, dp, old, labels, stream=s)
            ctx.mass_update(mcfg, n, k, labels, masses, stream=s)
            ctx.centroid_update(ccfg, d, n, k, dp, sums, labels, masses, stream=s)
            ctx.divide(d, k, sums, masses, new, stream=s)
        cen[0], cen[1] = new, old

    def restart():
        # first-x centroids, zeroed labels: the configuration's start
        cen[0].copy_from(init)
        ctx.fill_zero(labels)
        hip.synchronize()

    def hot_name(strategy):
        if args.pipeline != "single_stage":
            return {"unroll_vector_bounded": "LabelingUnrollVectorBounded"}.get(args.stage_labeling,
                                                                              "LabelingUnrollVector")
        return {"feature_sum": "FusedFeatureSum",
                "cluster_merge_bounded": "FusedClusterMergeBounded"}.get(strategy, "FusedClusterMerge")

    def measure(strategy):
        """W warm-up iterations, then K timed iterations from the start again."""
        fcfg = ctx.config(clkmeans.STAGE_FUSED, strategy)
        restart()
        for _ in range(args.warmup):
            iteration(fcfg)
        stream.synchronize()
        restart()
        # Timed region: K iterations, every kernel bracketed by HIP events on
        # the stream it runs on.
        ctx.clear_records()
        ctx.profiling(True)
        e0, e1 = hip.Event(), hip.Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for it in range(args.steps):
            ctx.set_iteration(it)
            iteration(fcfg)
        e1.record(stream)
        e1.synchronize()
        hip.synchronize()
        wall = time.perf_counter() - w0
        ctx.profiling(False)
        by_name = {}
        for name, _, ms in ctx.records():
            by_name.setdefault(name, []).append(ms)
        hot = hot_name(strategy)
        return {"cfg": fcfg, "ms_per_step": e0.elapsed_ms(e1) / args.steps, "wall": wall, "by_name": by_name,
                "hot": hot, "hot_ms": float(np.mean(by_name[hot])) if hot in by_name else float("nan"),
                "hot_each": by_name.get(hot, [])}

    compare = args.pipeline == "single_stage" and args.strategy != "cluster_merge" and not args.no_compare
    exhaustive = measure("cluster_merge") if compare else None
    run = measure(args.strategy)
    ms_per_step, wall, by_name, hot, hot_ms = (run["ms_per_step"], run["wall"], run["by_name"], run["hot"],
                                               run["hot_ms"])
    fcfg = run["cfg"]
    bytes_per_launch = n * (4 * d + 4) if args.pipeline == "single_stage" else n * (4 * d + 4)
    achieved = bytes_per_launch / (hot_ms * 1e-3) / 1e9
    log("kernels (avg ms): " + ", ".join("%s=%.4f x%d" % (kk, np.mean(v), len(v)) for kk, v in by_name.items()))

    # One more step, untimed: its labels for the parity spot check.
    c_in = cen[0].to_host()
    iteration(fcfg)
    stream.synchronize()
    lab = labels.to_host()
    mass_ok = bool(int(masses.to_host().astype(np.int64).sum()) == n)

    cpu = None
    parity = None
    idx = np.zeros(0, np.int64)
    if not args.no_cpu:
        # The CPU baseline leg: the oracle timed on the host, and as the
        # checker of a sample of the GPU labels.
        from oracle import oracle as O
        rng = np.random.default_rng(0)
        idx = np.sort(rng.choice(n, size=min(n, 1 << 16), replace=False))
        parity = bool(np.array_equal(lab[idx], O.label(np.ascontiguousarray(pts[:, idx]), c_in)))
        iters = max(1, args.cpu_iters)
        t = time.perf_counter()
        O.lloyd_mt(pts, init, iters, threads)
        dt = time.perf_counter() - t
        cpu = {"value": iters / dt, "unit": "iter/s", "cores": threads, "kind": "port",
               "sample": "%d full iterations of the same %d-point workload, fused cluster_merge "
                         "restatement (oracle/kmeans_oracle.c ora_lloyd_mt_f32), %d threads" % (iters, n, threads)}

    traffic = None
    pmc = os.path.join(REPO, "profiles", "fused_%s_n2p27_d4_k10_pmc.json" % args.strategy)
    if os.path.exists(pmc) and args.pipeline == "single_stage" and n == (1 << 27) and d == 4 and k == 10:
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    def summary(r, strategy):
        ach = bytes_per_launch / (r["hot_ms"] * 1e-3) / 1e9
        return {"strategy": strategy, "value": 1000.0 / r["ms_per_step"], "unit": "iter/s",
                "ms_per_step": r["ms_per_step"], "kernel": r["hot"], "kernel_ms": r["hot_ms"],
                "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS}

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_step,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain fused Lloyd, N=2^%d d=%d k=%d float32 feature-major, sigma=1 blobs "
                               "in [-50,50]^%d, seed 0, first-x init" % (args.log2n, d, k, d),
                   "pipeline": args.pipeline, "strategy": args.strategy if args.pipeline == "single_stage" else
                   "%s/part_private/cluster_merge" % args.stage_labeling, "n": n, "d": d, "k": k},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": hot,
                     "kernel_ms": hot_ms, "bytes_per_launch": bytes_per_launch,
                     "kernel_ms_iterations_1_to_3": [round(x, 4) for x in run["hot_each"][:3]],
                     "kernel_ms_last": round(run["hot_each"][-1], 4) if run["hot_each"] else None},
        "exhaustive": summary(exhaustive, "cluster_merge") if exhaustive else None,
        "cpu_baseline": cpu,
        "wall_s": wall,
        "parity": {"labels_sample": int(idx.size), "labels_equal": parity, "mass_sum_equal_n": mass_ok},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def dense(args):
    """The dense-distance configuration (d=64, k=1024) on the MFMA labeling."""
    import clkmeans
    from clkmeans import hip

    log2n = args.log2n if args.log2n != 27 else 24
    n, d, k = 1 << log2n, 64, 1024
    steps = args.steps if args.steps != 30 else 20
    threads = min(16, os.cpu_count() or 1)
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
    ctx.clear_records()
    ctx.profiling(True)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for it in range(steps):
        ctx.set_iteration(it)
        iteration(lcfg)
    e1.record(stream)
    e1.synchronize()
    ctx.profiling(False)
    ms_per_step = e0.elapsed_ms(e1) / steps
    by_name = {}
    for name, _, ms in ctx.records():
        by_name.setdefault(name, []).append(ms)
    log("kernels (avg ms): " + ", ".join("%s=%.4f x%d" % (kk, np.mean(v), len(v)) for kk, v in by_name.items()))
    hot = "LabelingMfmaF16" if args.labeling == "mfma_fp16" else \
          "LabelingMfmaBF16" if args.labeling == "mfma_bf16" else "LabelingGlobal"
    hot_ms = float(np.mean(by_name[hot])) if hot in by_name else float("nan")
    flops = 2.0 * n * k * d
    achieved = flops / (hot_ms * 1e-3) / 1e12
    refined = ctx.labeling_refined_points() if args.labeling.startswith("mfma") else None

    # one more labeling, untimed: its labels for the parity spot check
    c_in = cen[0].to_host()
    ctx.labeling(lcfg, d, n, k, dp, cen[0], labels, stream=s)
    stream.synchronize()
    lab = labels.to_host()
    parity = None
    idx = np.zeros(0, np.int64)

    # the plain FP32 labeling (unroll_vector) for comparison, one call
    plain_ms = None
    if not args.no_cpu:
        pcfg = ctx.config(clkmeans.STAGE_LABELING, "unroll_vector")
        ctx.clear_records()
        ctx.profiling(True)
        ctx.labeling(pcfg, d, n, k, dp, cen[0], labels, stream=s)
        stream.synchronize()
        ctx.profiling(False)
        plain_ms = float(sum(ms for _, _, ms in ctx.records()))
        plain_equal = bool(np.array_equal(labels.to_host(), lab))
    cpu = None
    if not args.no_cpu:
        # The CPU baseline leg: the oracle timed on the host, and as the
        # checker of a sample of the GPU labels.
        from oracle import oracle as O
        idx = np.sort(np.random.default_rng(0).choice(n, size=4096, replace=False))
        parity = bool(np.array_equal(lab[idx], O.label(np.ascontiguousarray(pts[:, idx]), c_in)))
        ns = 1 << 16
        t = time.perf_counter()
        O.lloyd_mt(np.ascontiguousarray(pts[:, :ns]), init, 1, threads)
        dt = (time.perf_counter() - t) * (n / ns)
        cpu = {"value": 1.0 / dt, "unit": "iter/s", "cores": threads, "kind": "port",
               "sample": "one fused cluster_merge iteration (oracle/kmeans_oracle.c ora_lloyd_mt_f32) over "
                         "the first 2^16 of the 2^%d points against the same 1024 centroids, %d threads, "
                         "scaled by N/2^16" % (log2n, threads)}
    out = {
        "metric": "Lloyd iter/s, dense distances (d=64, k=1024), MFMA labeling with exact refinement",
        "value": 1000.0 / ms_per_step,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-mfma+f32-exact" if args.labeling == "mfma_fp16" else
                 "bf16-mfma+f32-exact" if args.labeling == "mfma_bf16" else "f32",
        "data": "synthetic",
        "config": {"workload": "dense Lloyd, N=2^%d d=64 k=1024 float32 feature-major, sigma=0.5 blobs in "
                               "[-1,1]^64, seed 0, first-x init; three_stage: %s labeling, part_local masses, "
                               "feature-slice FP64 sums, FP64 divide" % (log2n, args.labeling),
                   "n": n, "d": d, "k": k},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA16_PEAK_TFLOPS, "traffic": None, "kernel": hot,
                     "kernel_ms": hot_ms, "flops_per_launch": flops},
        "cpu_baseline": cpu,
        "refined_points_last_call": refined,
        "plain_fp32_labeling_ms": plain_ms,
        "parity": {"labels_sample": int(idx.size), "labels_equal": parity,
                   "plain_fp32_labels_equal": plain_equal if plain_ms is not None else None},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def batch(args):
    """Many uneven jobs: the batch engine against the serial per-job loop."""
    import clkmeans
    from clkmeans import hip

    iters = 25
    mix = clkmeans.job_mix(args.jobs, seed=0)
    threads = min(16, os.cpu_count() or 1)
    ctx = clkmeans.Context(0)
    stream = hip.Stream()
    s = stream.handle
    probe = clkmeans.Batch(ctx, 1, 1 << 20, 32, 256)
    need = sum(probe.job_bytes(n, d, k) for n, d, k, _ in mix)
    probe.close()
    arena = int(need * 1.02) + (64 << 20)
    eng = clkmeans.Batch(ctx, len(mix), arena, 32, 256, labeling=args.batch_labeling)
    ser = clkmeans.Batch(ctx, len(mix), arena, 32, 256) if not args.no_serial else None
    t0 = time.time()
    slots = []
    checks = {}
    small = [j for j, (n, d, k, _) in enumerate(mix) if n * k * d <= 2e7]
    check_ids = set(np.random.default_rng(1).choice(small, size=min(12, len(small)), replace=False).tolist())
    for j, (n, d, k, seed) in enumerate(mix):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed, threads=threads)
        init = np.ascontiguousarray(pts[:, :k])
        slots.append((eng.add(pts, init, iters, stream=s), ser.add(pts, init, iters, stream=s) if ser else None))
        if j in check_ids:
            checks[j] = (pts, init)
    log("generated and loaded %d jobs (%.2f GB of points) in %.1fs" %
        (len(mix), sum(n * d * 4 for n, d, _, _ in mix) / 1e9, time.time() - t0))
    points = sum(n for n, _, _, _ in mix)
    bytes_per_iter = sum(n * (4 * d + 4) for n, d, _, _ in mix)

    # tile sizes by the autotuner (dry passes over these jobs), then the
    # batch engine: graph capture on the first step, 25 replays timed
    tuned = eng.autotune(s)
    log("autotuned tile_work = %d" % tuned)
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
    states, active = eng.poll(s)
    # serial per-job loop on the same GPU
    stream.synchronize()
    w0 = time.perf_counter()
    if ser:
        ser.run_serial(s)
    stream.synchronize()
    serial_ms = (time.perf_counter() - w0) * 1e3 if ser else None

    # parity: every job batch == serial (labels, masses exact; centroids), a
    # sample against the oracle
    # centroid differences are measured against the job's largest centroid
    # coordinate (coordinates near 0 would inflate a per-element ratio)
    lab_equal, mass_equal, worst = 0, 0, 0.0
    sample_out = {}
    for j, (a, b_) in enumerate(slots):
        ca, la, ma, ita, _ = eng.take(a, s)
        if j in checks:
            sample_out[j] = (ca, la, ma)
        if not ser:
            continue
        cb, lb, mb, itb, _ = ser.take(b_, s)
        lab_equal += int(bool(np.array_equal(la, lb)) and ita == itb == iters)
        mass_equal += int(np.array_equal(ma, mb))
        live = mb > 0
        if live.any():
            scale = max(float(np.abs(cb[:, live]).max()), 1e-30)
            worst = max(worst, float(np.abs(ca[:, live].astype(np.float64) - cb[:, live]).max()) / scale)
    eng.close()
    if ser:
        ser.close()

    cpu = None
    oracle_ok, fp64_err = None, None
    if not args.no_cpu:
        # The CPU baseline leg.  The oracle as the checker of the sampled
        # jobs: labels and masses equal to its Lloyd loop, centroids against
        # its FP64 sums (relative to the cluster's mean |x| where coordinates
        # are near 0, as in the tests).
        from oracle import oracle as O
        oracle_ok, fp64_err = True, 0.0
        for j, (ca, la, ma) in sample_out.items():
            pts, init = checks[j]
            rc, rl, rm = O.lloyd(pts, init, iters)
            oracle_ok &= bool(np.array_equal(la, rl)) and bool(np.array_equal(ma, rm))
            live = rm > 0
            ref64 = O.sums_f64(pts, rl, init.shape[1])[:, live] / rm[live].astype(np.float64)
            absx = np.stack([np.bincount(rl, weights=np.abs(pts[f].astype(np.float64)),
                                         minlength=init.shape[1])[live] for f in range(pts.shape[0])])
            scale = np.maximum(np.abs(ref64), absx / rm[live])
            fp64_err = max(fp64_err, float((np.abs(ca[:, live] - ref64) / np.maximum(scale, 1e-30)).max()))
        # oracle (multi-threaded fused restatement) on the first 64 jobs,
        # scaled by total work n*k*d
        sub = mix[:64]
        t = time.perf_counter()
        for n, d, k, seed in sub:
            pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed, threads=threads)
            O.lloyd_mt(pts, np.ascontiguousarray(pts[:, :k]), iters, threads)
        dt = time.perf_counter() - t
        w_sub = sum(n * k * d for n, d, k, _ in sub)
        w_all = sum(n * k * d for n, d, k, _ in mix)
        cpu = {"value": len(mix) / (dt * w_all / w_sub), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "the first 64 jobs, 25 iterations each, fused cluster_merge restatement "
                         "(oracle/kmeans_oracle.c ora_lloyd_mt_f32, %d threads, includes data generation), "
                         "scaled by total n*k*d work" % threads}
    achieved = bytes_per_iter / (np.mean(step_ms) * 1e-3) / 1e9 if step_ms else None
    out = {
        "metric": "batched vs plain per-job solves/s (4096 uneven jobs, 25 iterations)",
        "value": len(mix) / (batch_ms * 1e-3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": iters,
        "warmup": 0,
        "ms_per_step": batch_ms / iters,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.batch_labeling == "unroll_vector" else
                 args.batch_labeling.split("_")[1] + "-mfma+f32-exact",
        "data": "synthetic",
        "config": {"workload": "%d jobs: N log-uniform [2^10,2^18], d in {3,8,16,32}, k log-uniform [4,256] "
                               "capped at N/16, sigma=1 blobs in [-10,10]^d, per-job seeds, first-x init, "
                               "%d iterations; batch engine (one persistent launch per iteration, HIP graph), "
                               "%s labeling" % (len(mix), iters, args.batch_labeling),
                   "jobs": len(mix), "points": points, "labeling": args.batch_labeling,
                   "tile_work": tuned},
        "serial_per_job_loop": {"solves_per_s": len(mix) / (serial_ms * 1e-3), "total_ms": serial_ms}
                               if serial_ms else None,
        "speedup_vs_serial": serial_ms / batch_ms if serial_ms else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "kernel": "one batch iteration (size-class tile kernels + merge kernel)",
                     "bytes_per_launch": bytes_per_iter,
                     "kernel_ms": float(np.mean(step_ms)) if step_ms else None},
        "cpu_baseline": cpu,
        "parity": {"jobs_labels_equal_serial": lab_equal, "jobs_masses_equal_serial": mass_equal,
                   "max_centroid_diff_vs_serial_rel_job_scale": worst, "oracle_sample_jobs": len(checks),
                   "oracle_sample_equal": oracle_ok, "oracle_sample_centroid_rel_err_vs_fp64": fp64_err, "jobs_done": int((states == clkmeans.JOB_DONE).sum()),
                   "still_active": int(active)},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def serve(args):
    """Poisson arrivals of the batched mix plus 2% large jobs, continuous
    batching (clk_serve_run) against the serial plain-FP32 per-job loop
    (clk_serve_serial) on the same arrivals; first with the points already in
    HBM (value), then copied in from pinned host memory on admission."""
    import clkmeans
    from clkmeans import hip

    threads = min(16, os.cpu_count() or 1)
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
    arrivals = clkmeans.poisson_arrivals(count, args.rate, seed=3)
    log("generated %d jobs (%d large, %.2f GB of points) in %.1fs" %
        (count, len(large), sum(n * d * 4 for n, d, _ in specs) / 1e9, time.time() - t0))
    probe = clkmeans.Batch(ctx, 1, 1 << 20, 32, 256)
    big_bytes = probe.job_bytes(1 << 24, 32, 256) if large else probe.job_bytes(1 << 18, 32, 256)
    probe.close()

    def requests(resident):
        return [clkmeans.ServeRequest(dev[j] if resident else host[j].array, inits[j], arrivals[j],
                                      max_iterations=50, tolerance=1e-4, want_labels=False)
                for j in range(count)]

    def run(resident):
        reqs = requests(resident)
        eng = clkmeans.Batch(ctx, 1024, int(big_bytes * 12 + (8 << 30)), 32, 256, tile_work=args.tile_work,
                             labeling=args.batch_labeling)
        st = clkmeans.serve(eng, reqs, stream.handle)
        eng.close()
        sreqs = requests(resident)
        ss = clkmeans.serve_serial(ctx, sreqs, stream.handle)
        same_iters = sum(int(a.iterations == b.iterations) for a, b in zip(reqs, sreqs))
        same_mass = sum(int(np.array_equal(a.out_masses, b.out_masses)) for a, b in zip(reqs, sreqs))
        lat_large = [reqs[j].latency for j in large] or [0.0]
        return {
            "served": {"solves_per_s": st["solves_per_s"], "makespan_s": st["makespan"], "p50_s": st["p50"],
                       "p99_s": st["p99"], "mean_latency_s": st["mean_latency"],
                       "iterations_launched": int(st["steps"]), "max_active": int(st["max_active"]),
                       "large_job_latency_s": {"min": min(lat_large), "max": max(lat_large)}},
            "serial_per_job_loop": {"solves_per_s": ss["solves_per_s"], "makespan_s": ss["makespan"],
                                    "p50_s": ss["p50"], "p99_s": ss["p99"]},
            "speedup_vs_serial": st["solves_per_s"] / ss["solves_per_s"],
            "parity": {"jobs_same_iterations_as_serial": same_iters, "jobs_same_masses_as_serial": same_mass},
        }

    res = run(True)
    pcie = run(False)
    sr = res["served"]
    out = {
        "metric": "served vs plain per-job solves/s at tol 1e-4 (Poisson arrivals), p50/p99 latency",
        "value": sr["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": sr["iterations_launched"],
        "warmup": 0,
        "ms_per_step": sr["makespan_s"] * 1e3 / max(1, sr["iterations_launched"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.batch_labeling == "unroll_vector" else
                 args.batch_labeling.split("_")[1] + "-mfma+f32-exact",
        "data": "synthetic",
        "config": {"workload": "%d jobs, Poisson arrivals at %.0f jobs/s: the batched mix (N log-uniform "
                               "[2^10,2^18], d in {3,8,16,32}, k log-uniform [4,256] capped at N/16) with %d of "
                               "them large jobs (N=2^24, d=32, k=256; 4 distinct data sets); each until max "
                               "shift < 1e-4 or 50 iterations; points already in HBM (read in place)"
                               % (count, args.rate, len(large)),
                   "jobs": count, "rate": args.rate, "labeling": args.batch_labeling},
        **res,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
