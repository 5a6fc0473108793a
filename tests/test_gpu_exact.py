"""Exact centroid sums (the sum grid of clkmeans.h): every path, one result.

Each coordinate is rounded to nearest-even on a per-feature power-of-two grid
and the grid integers are added exactly, so:
  - the grid the library computes equals the oracle's (ora_grid_scales,
    pinned by its definition in tests/test_oracle.py);
  - every sum-accumulating op (fused strategies, centroid-update strategies,
    the large-k*d slice and global-atomic kernels) returns the oracle's exact
    grid sums bit for bit, and lloyd_step the centroids divided from them;
  - chunked accumulation under a pinned grid equals one pass;
  - the batch engine at any tile_work, the plain per-job loop
    (clk_lloyd_step) and the oracle's grid Lloyd loop agree bit for bit over
    25 iterations of overlapping blobs, with FP32 and MFMA labeling.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

FUSED = clkmeans.STAGE_FUSED
CENTROID = clkmeans.STAGE_CENTROID_UPDATE


def oracle_sums(pts, labels, k):
    return O.sums_grid(pts, labels, k, O.grid_scales(pts))


def test_grid_units_match_oracle(gpu_ctx):
    rng = np.random.default_rng(50)
    cases = [clkmeans.generate_blobs(10000, 4, 10, seed=0)[0]]
    for e in (-120, -20, 0, 30, 110):
        cases.append((rng.standard_normal((3, 777)) * 2.0 ** e).astype(np.float32))
    z = np.zeros((3, 64), np.float32)
    z[1, 5] = 1e-44
    z[2, 9] = np.nan
    z[2, 10] = 2.0
    z[2, 11] = -np.inf
    cases.append(z)
    for pts in cases:
        d, n = pts.shape
        units = gpu_ctx.sum_grid_units(d, n, dev(pts))
        np.testing.assert_array_equal(units, 1.0 / O.grid_scales(pts).astype(np.float64))


SHAPES = [  # n, d, k
    (65536, 4, 10), (50001, 3, 7), (40000, 1, 5), (30000, 8, 64), (20000, 17, 33),
    (12345, 32, 100), (9000, 64, 12), (20000, 32, 256), (8192, 2, 1),
]


@pytest.mark.parametrize("strategy", ["cluster_merge", "cluster_merge_bounded", "feature_sum"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "n%d_d%d_k%d" % s)
def test_fused_sums_bit_exact(gpu_ctx, strategy, shape):
    n, d, k = shape
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=2.0, box=10.0, seed=n + d + k)
    init = O.first_x(pts, k)
    dp, dc = dev(pts), dev(init)
    sums, masses = zeros((d, k), np.float64), zeros(k, np.uint32)
    labels = zeros(n, np.uint32)
    cfg = gpu_ctx.config(FUSED, strategy)
    gpu_ctx.fused(cfg, d, n, k, dp, dc, sums, labels, masses)
    hip.synchronize()
    lab = labels.to_host()
    np.testing.assert_array_equal(lab, O.label(pts, init))
    np.testing.assert_array_equal(masses.to_host(), O.mass(lab, k))
    ref = oracle_sums(pts, lab, k)
    np.testing.assert_array_equal(sums.to_host(), ref)
    # the whole iteration: centroids divided from the same sums
    newc = zeros((d, k), np.float32)
    gpu_ctx.lloyd_step(cfg, d, n, k, dp, dc, newc, sums, labels, masses)
    hip.synchronize()
    np.testing.assert_array_equal(newc.to_host(), O.divide_from_f64(ref, O.mass(lab, k)))


@pytest.mark.parametrize("strategy,kw", [("cluster_merge", {}), ("feature_sum", {}),
                                         ("feature_sum_pardim", {"local_features": 2, "thread_features": 2}),
                                         ("feature_sum_pardim", {})])
@pytest.mark.parametrize("shape", [(30000, 4, 10), (20001, 16, 40), (16384, 64, 1024), (4096, 64, 3000),
                                   (7000, 3, 9000)], ids=lambda s: "n%d_d%d_k%d" % s)
def test_centroid_update_strategies_bit_exact(gpu_ctx, strategy, kw, shape):
    # (16384, 64, 1024): the feature-slice kernel; (7000, 3, 9000) and
    # (4096, 64, 3000): the global-atomic pass (of grid values)
    n, d, k = shape
    rng = np.random.default_rng(n + k)
    pts = (rng.standard_normal((d, n)) * 7).astype(np.float32)
    lab = rng.integers(0, k, n).astype(np.uint32)
    sums = zeros((d, k), np.float64)
    cfg = gpu_ctx.config(CENTROID, strategy, **kw)
    gpu_ctx.centroid_update(cfg, d, n, k, dev(pts), sums, dev(lab))
    hip.synchronize()
    np.testing.assert_array_equal(sums.to_host(), oracle_sums(pts, lab, k))


def test_pinned_grid_chunks_equal_one_pass(gpu_ctx):
    """The _buffered form: the grid of all points pinned, the fused op called
    per chunk (accumulating) -- the same sums as one pass."""
    n, d, k = 100000, 5, 12
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=3.0, box=10.0, seed=9)
    init = O.first_x(pts, k)
    dc = dev(init)
    cfg = gpu_ctx.config(FUSED, "cluster_merge")
    one = zeros((d, k), np.float64)
    gpu_ctx.fused(cfg, d, n, k, dev(pts), dc, one, zeros(n, np.uint32), zeros(k, np.uint32))
    gpu_ctx.sum_grid_from_max(d, n, np.abs(pts).max(axis=1))
    try:
        acc = zeros((d, k), np.float64)
        masses = zeros(k, np.uint32)
        for c0 in range(0, n, 12288):
            chunk = np.ascontiguousarray(pts[:, c0:c0 + 12288])
            m = chunk.shape[1]
            gpu_ctx.fused(cfg, d, m, k, dev(chunk), dc, acc, zeros(m, np.uint32), masses)
        hip.synchronize()
    finally:
        gpu_ctx.sum_grid_clear()
    np.testing.assert_array_equal(acc.to_host(), one.to_host())
    np.testing.assert_array_equal(acc.to_host(), oracle_sums(pts, O.label(pts, init), k))


def test_nonfinite_and_off_grid_points(gpu_ctx):
    n, d, k = 4096, 3, 4
    pts, _ = clkmeans.generate_blobs(n, d, k, seed=5)
    init = O.first_x(pts, k)
    pts[0, 100] = np.nan      # label 0 (NaN distances): cluster 0's feature 0 sum is NaN
    pts[1, 200] = np.inf
    cfg = gpu_ctx.config(FUSED, "cluster_merge")
    sums = zeros((d, k), np.float64)
    labels = zeros(n, np.uint32)
    gpu_ctx.fused(cfg, d, n, k, dev(pts), dev(init), sums, labels, zeros(k, np.uint32))
    hip.synchronize()
    lab = labels.to_host()
    np.testing.assert_array_equal(lab, O.label(pts, init))
    np.testing.assert_array_equal(sums.to_host(), oracle_sums(pts, lab, k))
    s = sums.to_host()
    assert np.isnan(s[0, lab[100]]) and (np.isinf(s[1, lab[200]]) or np.isnan(s[1, lab[200]]))
    # a grid pinned for data 1000x smaller: the values are off the grid (NaN)
    gpu_ctx.sum_grid_from_max(d, n, np.abs(pts[:, :64]).max(axis=1) * 1e-3)
    try:
        good = np.nan_to_num(pts, nan=0.0, posinf=0.0)
        s2 = zeros((d, k), np.float64)
        gpu_ctx.fused(cfg, d, n, k, dev(good), dev(init), s2, zeros(n, np.uint32), zeros(k, np.uint32))
        hip.synchronize()
    finally:
        gpu_ctx.sum_grid_clear()
    assert np.isnan(s2.to_host()).any()


# ---------------------------------------------------------------------------
# Batch engine == plain per-job loop == oracle, bit for bit

OVERLAP = [  # n, d, k (sigma 3 blobs in [-10, 10]^d: many points near cell boundaries)
    (20000, 4, 10), (8191, 3, 25), (30000, 8, 64), (12000, 16, 40), (5000, 32, 100), (40000, 32, 256),
    (3000, 8, 150),
]


def plain_loop(ctx, pts, init, iters):
    """clk_lloyd_step per iteration on the job's own buffers (fused
    cluster_merge), with the job's grid pinned."""
    d, n = pts.shape
    k = init.shape[1]
    dp = dev(pts)
    cen = [dev(init), zeros((d, k), np.float32)]
    sums, labels, masses = zeros((d, k), np.float64), zeros(n, np.uint32), zeros(k, np.uint32)
    cfg = ctx.config(FUSED, "cluster_merge")
    ctx.sum_grid_from_points(d, n, dp)
    try:
        for it in range(iters):
            ctx.lloyd_step(cfg, d, n, k, dp, cen[it % 2], cen[(it + 1) % 2], sums, labels, masses)
        hip.synchronize()
    finally:
        ctx.sum_grid_clear()
    return cen[iters % 2].to_host(), labels.to_host(), masses.to_host()


@pytest.fixture(scope="module")
def overlap_jobs():
    jobs = []
    for i, (n, d, k) in enumerate(OVERLAP):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=3.0, box=10.0, seed=700 + i)
        jobs.append((pts, O.first_x(pts, k)))
    return jobs


@pytest.fixture(scope="module")
def plain_results(gpu_ctx, overlap_jobs):
    out = []
    for pts, init in overlap_jobs:
        got = plain_loop(gpu_ctx, pts, init, 25)
        rc, rl, rm = O.lloyd(pts, init, 25, mode=O.MODE_GRID)
        np.testing.assert_array_equal(got[1], rl)
        np.testing.assert_array_equal(got[2], rm)
        np.testing.assert_array_equal(got[0], rc)
        out.append(got)
    return out


@pytest.mark.parametrize("labeling", ["unroll_vector", "mfma_bf16"])
@pytest.mark.parametrize("tile_work", [4, 64])
def test_batch_equals_plain_loop_bitwise(gpu_ctx, overlap_jobs, plain_results, labeling, tile_work):
    s = hip.Stream()
    b = clkmeans.Batch(gpu_ctx, max_jobs=16, arena_bytes=512 << 20, max_d=32, max_k=256,
                       tile_work=tile_work, labeling=labeling)
    try:
        slots = [b.add(pts, init, max_iterations=25, stream=s.handle) for pts, init in overlap_jobs]
        for _ in range(25):
            b.step(stream=s.handle)
        s.synchronize()
        for slot, (pc, pl, pm) in zip(slots, plain_results):
            c, lab, m, it, _ = b.take(slot, stream=s.handle)
            assert it == 25
            np.testing.assert_array_equal(lab, pl)
            np.testing.assert_array_equal(m, pm)
            np.testing.assert_array_equal(c, pc)
    finally:
        b.close()
        s.synchronize()


def test_batch_run_serial_equals_batch(gpu_ctx, overlap_jobs):
    """clk_batch_run_serial (the plain loop on the batch's own buffers) and
    clk_batch_step give the same bits, for every job."""
    s = hip.Stream()
    res = []
    for serial in (False, True):
        b = clkmeans.Batch(gpu_ctx, max_jobs=16, arena_bytes=512 << 20, max_d=32, max_k=256,
                           labeling="mfma_fp16")
        try:
            slots = [b.add(pts, init, max_iterations=12, stream=s.handle) for pts, init in overlap_jobs]
            if serial:
                b.run_serial(stream=s.handle)
            else:
                for _ in range(12):
                    b.step(stream=s.handle)
            s.synchronize()
            res.append([b.take(slot, stream=s.handle) for slot in slots])
        finally:
            b.close()
    for a, c in zip(*res):
        for x, y in zip(a[:4], c[:4]):
            np.testing.assert_array_equal(x, y)
