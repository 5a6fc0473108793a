"""The double / uint64 / uint64 type triple (bench.cpp:469-503) on the GPU,
against the oracle's FP64 functions: labels and masses exact, FP64 sums and
centroids within a bound of FP64 rounding for a different summation order.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

FUSED = clkmeans.STAGE_FUSED
LAB = clkmeans.STAGE_LABELING
MASS = clkmeans.STAGE_MASS_UPDATE
CU = clkmeans.STAGE_CENTROID_UPDATE


def blobs64(n, d, k, seed, sigma=1.0, box=20.0):
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=seed)
    rng = np.random.default_rng(seed)
    # below float32 resolution too: a float64 path that rounds through
    # float32 would mislabel points
    return pts.astype(np.float64) + rng.standard_normal(pts.shape) * 1e-9


def sums_close(got, points, labels, k):
    ref = O.sums_f64_dp(points, labels, k)
    absx = np.zeros_like(ref)
    for f in range(points.shape[0]):
        absx[f] = np.bincount(labels.astype(np.int64), weights=np.abs(points[f]), minlength=k)
    assert np.all(np.abs(got - ref) <= 1e-13 * absx + 1e-300)


CASES = [
    # n, d, k, config overrides
    (20000, 4, 10, {}),
    (7777, 3, 5, {"local_size": 64}),
    (5000, 1, 3, {}),
    (4096, 2, 1, {}),
    (3000, 8, 37, {"local_size": 1024}),
    (3000, 16, 20, {"global_size": 4096, "local_size": 128}),
    (2048, 32, 12, {}),
    (1500, 33, 6, {}),          # beyond the register buckets: global path
    (2000, 16, 1200, {}),       # centroids beyond LDS: global path
    (7, 4, 10, {}),             # n < k
]


@pytest.mark.parametrize("n,d,k,kw", CASES)
def test_f64_operators_match_oracle(gpu_ctx, n, d, k, kw):
    pts = blobs64(n, d, k, seed=n + d)
    cen = np.ascontiguousarray(pts[:, np.arange(k) % n])
    ref = O.label_f64(pts, cen)
    dp, dc = dev(pts), dev(cen)

    dl = zeros((n,), np.uint64)
    gpu_ctx.labeling(gpu_ctx.config(LAB, "unroll_vector", **kw), d, n, k, dp, dc, dl)
    hip.synchronize()
    np.testing.assert_array_equal(dl.to_host(), ref)

    m0 = np.arange(k, dtype=np.uint64) * 5  # accumulate semantics
    dm = dev(m0)
    gpu_ctx.mass_update(gpu_ctx.config(MASS, "part_private", **kw), n, k, dl, dm)
    hip.synchronize()
    np.testing.assert_array_equal(dm.to_host() - m0, O.mass_u64(ref, k))

    s0 = np.full((d, k), 0.25)
    ds = dev(s0)
    gpu_ctx.centroid_update(gpu_ctx.config(CU, "cluster_merge", **kw), d, n, k, dp, ds, dl)
    hip.synchronize()
    sums_close(ds.to_host() - s0, pts, ref, k)

    # fused: labels overwritten, masses and sums accumulated
    dl2 = dev(np.full(n, 99, np.uint64))
    dm2, ds2 = dev(m0), dev(s0)
    gpu_ctx.fused(gpu_ctx.config(FUSED, "cluster_merge", **kw), d, n, k, dp, dc, ds2, dl2, dm2)
    hip.synchronize()
    np.testing.assert_array_equal(dl2.to_host(), ref)
    np.testing.assert_array_equal(dm2.to_host() - m0, O.mass_u64(ref, k))
    sums_close(ds2.to_host() - s0, pts, ref, k)


@pytest.mark.parametrize("strategy", ["cluster_merge", "feature_sum", "cluster_merge_bounded"])
def test_f64_lloyd_matches_oracle_loop(gpu_ctx, strategy):
    n, d, k, iters = 30000, 5, 12, 6
    pts = blobs64(n, d, k, seed=3, sigma=2.0, box=10.0)
    init = np.ascontiguousarray(pts[:, :k])
    cfg = gpu_ctx.config(FUSED, strategy)
    dp = dev(pts)
    cen = [dev(init), zeros((d, k), np.float64)]
    sums, masses, labels = zeros((d, k), np.float64), zeros((k,), np.uint64), zeros((n,), np.uint64)
    shift = zeros((1,), np.float64)
    c_host = init.copy()
    for it in range(iters):
        gpu_ctx.lloyd_step(cfg, d, n, k, dp, cen[0], cen[1], sums, labels, masses, max_shift=shift)
        hip.synchronize()
        rl = O.label_f64(pts, c_host)
        np.testing.assert_array_equal(labels.to_host(), rl)
        rm = O.mass_u64(rl, k)
        np.testing.assert_array_equal(masses.to_host(), rm)
        new = cen[1].to_host()
        with np.errstate(invalid="ignore", divide="ignore"):
            ref = O.sums_f64_dp(pts, rl, k) / rm.astype(np.float64)
        live = rm != 0
        np.testing.assert_allclose(new[:, live], ref[:, live], rtol=1e-12, atol=1e-12)
        assert np.all(np.isnan(new[:, ~live]))
        moved = np.sqrt(((new[:, live] - c_host[:, live]) ** 2).sum(0)).max() if live.any() else 0.0
        assert abs(shift.to_host()[0] - moved) <= 1e-9 * max(1.0, moved)
        c_host = new
        cen = [cen[1], cen[0]]


def test_f64_ties_nan_and_range(gpu_ctx):
    # equidistant centroids: the lower index; a NaN centroid never wins; all
    # distances NaN: label 0; labels >= k are skipped by the mass update
    pts = np.array([[0.0, 1.0, 5.0, 2.0]])
    cen = np.array([[2.0, 0.0, 2.0, np.nan]])
    dl = zeros((4,), np.uint64)
    gpu_ctx.labeling(gpu_ctx.config(LAB, "unroll_vector"), 1, 4, 4, dev(pts), dev(cen), dl)
    hip.synchronize()
    np.testing.assert_array_equal(dl.to_host(), O.label_f64(pts, cen))
    np.testing.assert_array_equal(dl.to_host(), [1, 0, 0, 0])
    allnan = np.full((1, 3), np.nan)
    gpu_ctx.labeling(gpu_ctx.config(LAB, "unroll_vector"), 1, 4, 3, dev(pts), dev(allnan), dl)
    hip.synchronize()
    np.testing.assert_array_equal(dl.to_host(), [0, 0, 0, 0])
    lab = np.array([0, 1, 7, 2**40, 1], np.uint64)
    dm = zeros((3,), np.uint64)
    gpu_ctx.mass_update(gpu_ctx.config(MASS, "global_atomic"), 5, 3, dev(lab), dm)
    hip.synchronize()
    np.testing.assert_array_equal(dm.to_host(), [1, 2, 0])
