"""MFMA (16-bit operand) labeling with the exact FP32 relabel: labels must be
bit-identical to the oracle's sequential-fma labeling (kmeans_naive.cpp:49-63,
lloyd_fused_feature_sum.cl:210-239) for both operand precisions -- on
well-separated blobs, on data built to produce exact and near ties, and on
special values that force the exact scan.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

LAB = clkmeans.STAGE_LABELING
PRECISIONS = ["mfma_bf16", "mfma_fp16"]


def mfma_labels(ctx, strategy, pts, cen):
    d, n = pts.shape
    k = cen.shape[1]
    dl = zeros(max(n, 1), np.uint32)
    ctx.labeling(ctx.config(LAB, strategy), d, n, k, dev(pts), dev(cen), dl)
    hip.synchronize()
    return dl.to_host()[:n]


def check(ctx, strategy, pts, cen):
    got = mfma_labels(ctx, strategy, pts, cen)
    ref = O.label(np.ascontiguousarray(pts), np.ascontiguousarray(cen))
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, "%d labels differ, first at point %d: got %d want %d" % (
        bad.size, bad[0], got[bad[0]], ref[bad[0]])
    return got


BLOB_CASES = [
    # n, d, k, sigma, box
    ((1 << 14) + 5, 64, 1024, 0.5, 1.0),   # the dense-distance shape (d=64, k=1024), reduced n
    (20000, 4, 10, 1.0, 50.0),
    (7777, 17, 100, 1.0, 10.0),           # partial feature block, partial tile
    (5000, 32, 256, 1.0, 10.0),
    (3000, 100, 50, 0.5, 1.0),
    (3001, 128, 300, 0.5, 1.0),
    (1000, 1, 3, 1.0, 10.0),
    (65, 2, 1, 1.0, 10.0),                # k == 1
    (63, 3, 17, 1.0, 10.0),               # one partial wave
]


@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("n,d,k,sigma,box", BLOB_CASES)
def test_mfma_blobs_match_oracle(gpu_ctx, strategy, n, d, k, sigma, box):
    pts, centres = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=n + d + k)
    # first_x-like centroids (data points) and perturbed blob centres
    check(gpu_ctx, strategy, pts, np.ascontiguousarray(pts[:, :k]) if n >= k else centres)
    rng = np.random.default_rng(k)
    check(gpu_ctx, strategy, pts, (centres + rng.standard_normal(centres.shape) * sigma).astype(np.float32))


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_uniform_high_dim_near_ties(gpu_ctx, strategy):
    # Uniform points against uniform centroids in 64 dimensions: distances
    # concentrate, best-to-second margins fall inside the 16-bit error bound
    # for many points, and most waves take the exact rescan.
    rng = np.random.default_rng(7)
    pts = rng.random((64, 4096), dtype=np.float32)
    cen = rng.random((64, 1024), dtype=np.float32)
    check(gpu_ctx, strategy, pts, cen)
    assert gpu_ctx.labeling_refined_points() > 0


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_exact_ties_lowest_index(gpu_ctx, strategy):
    # Integer lattice: centroids at even coordinates, points at every lattice
    # point -> exact FP32 ties between up to 2^d centroids.
    d = 3
    g = np.arange(0, 9, 2, dtype=np.float32)
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), 0).reshape(d, -1)
    cen = np.ascontiguousarray(cen[:, np.random.default_rng(1).permutation(cen.shape[1])])
    p = np.arange(-1, 10, dtype=np.float32)
    pts = np.ascontiguousarray(np.stack(np.meshgrid(p, p, p, indexing="ij"), 0).reshape(d, -1))
    got = check(gpu_ctx, strategy, pts, cen)
    assert got.max() < cen.shape[1]


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_duplicate_centroids(gpu_ctx, strategy):
    pts, centres = clkmeans.generate_blobs(5000, 16, 40, sigma=1.0, box=10.0, seed=5)
    cen = np.ascontiguousarray(np.concatenate([centres, centres[:, ::-1], centres], axis=1))
    got = check(gpu_ctx, strategy, pts, cen)
    assert got.max() < 40 * 2  # the third copy never wins (lower indices tie first)


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_special_values(gpu_ctx, strategy):
    pts, centres = clkmeans.generate_blobs(4096, 8, 20, sigma=1.0, box=10.0, seed=9)
    cen = centres.copy()
    cen[:, 3] = np.nan                  # NaN centroid never wins
    cen[2, 7] = np.inf                  # inf coordinate never wins
    cen[0, 11] = -np.inf
    check(gpu_ctx, strategy, pts, cen)
    # every centroid NaN -> label 0 everywhere
    got = check(gpu_ctx, strategy, pts, np.full_like(cen, np.nan))
    assert not got.any()


@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("scale", [1e-6, 1e-30, 4e4, 1e5, 1e18])
def test_mfma_scales(gpu_ctx, strategy, scale):
    # Tiny values (16-bit subnormals / flushed), values whose -2c overflows
    # FP16 (exact scan for the whole call), norms that overflow the bound.
    pts, centres = clkmeans.generate_blobs(3000, 12, 30, sigma=1.0, box=10.0, seed=11)
    check(gpu_ctx, strategy, (pts * np.float32(scale)).astype(np.float32),
          (centres * np.float32(scale)).astype(np.float32))


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_mixed_scales_points(gpu_ctx, strategy):
    pts, centres = clkmeans.generate_blobs(4096, 6, 12, sigma=1.0, box=10.0, seed=13)
    pts = pts.copy()
    pts[:, 100:164] *= np.float32(1e19)     # one wave of huge points (exact scan)
    pts[:, 500] = np.nan                     # NaN point -> label 0
    pts[:, 501] = np.inf
    check(gpu_ctx, strategy, pts, centres)


@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000])
def test_mfma_small_n(gpu_ctx, strategy, n):
    rng = np.random.default_rng(n)
    pts = rng.standard_normal((5, n)).astype(np.float32)
    cen = rng.standard_normal((5, 33)).astype(np.float32)
    check(gpu_ctx, strategy, pts, cen)


@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("d,k", [(129, 10), (64, 2000)])
def test_mfma_unsupported_sizes_fall_back(gpu_ctx, strategy, d, k):
    # d > 128, or a 16-bit table larger than LDS: the unroll_vector kernel runs.
    rng = np.random.default_rng(d + k)
    pts = rng.standard_normal((d, 700)).astype(np.float32)
    cen = rng.standard_normal((d, k)).astype(np.float32)
    check(gpu_ctx, strategy, pts, cen)


@pytest.mark.parametrize("strategy,frac", [("mfma_bf16", 0.10), ("mfma_fp16", 0.02)])
def test_mfma_refinement_is_the_exception_on_blobs(gpu_ctx, strategy, frac):
    # d=64, k=1024 blobs (sigma 0.5, centres in [-1,1]^64), first-k init:
    # the margin test settles all but a small fraction of points.
    n, d, k = 1 << 16, 64, 1024
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=0.5, box=1.0, seed=0)
    cen = np.ascontiguousarray(pts[:, :k])
    got = mfma_labels(gpu_ctx, strategy, pts, cen)
    refined = gpu_ctx.labeling_refined_points()
    assert 0 < refined < frac * n, refined
    sample = np.arange(0, n, 7)
    np.testing.assert_array_equal(got[sample], O.label(np.ascontiguousarray(pts[:, sample]), cen))



@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("n,d,k", [((1 << 20) + 77, 32, 256), ((1 << 19) + 3, 64, 512)])
def test_mfma_many_batches_per_wave_match_fp32_scan(gpu_ctx, strategy, n, d, k):
    # Enough points that every wave of the persistent streaming kernel labels
    # several 64-point batches (the grid holds one work-group of 16 waves per
    # CU); the labels must equal the exact FP32 scan's (unroll_vector, itself
    # pinned to the oracle by the other tests) on every point.
    pts, centres = clkmeans.generate_blobs(n, d, k, sigma=0.5, box=1.0, seed=d + k)
    cen = np.ascontiguousarray(pts[:, :k])
    got = mfma_labels(gpu_ctx, strategy, pts, cen)
    ref = mfma_labels(gpu_ctx, "unroll_vector", pts, cen)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, "%d labels differ, first at point %d" % (bad.size, bad[0])


# ---- the pinned 16-bit operand copy (clk_labeling_pin_points) ------------

PIN_CASES = [
    # n, d, k, sigma, box
    ((1 << 14) + 5, 64, 1024, 0.5, 1.0),  # the dense shape, reduced n, partial last step
    (20000, 40, 64, 1.0, 10.0),           # partial second K-block
    (7777, 17, 100, 1.0, 10.0),
    (3001, 128, 300, 0.5, 1.0),           # four K-blocks
    (63, 3, 17, 1.0, 10.0),               # one partial step
]


def pinned_labels(ctx, strategy, dp, n, d, cen, pin_strategy=None, pin_n=None):
    k = cen.shape[1]
    dl = zeros(max(n, 1), np.uint32)
    ctx.labeling_pin_points(pin_strategy or strategy, d, pin_n or n, dp)
    ctx.labeling(ctx.config(LAB, strategy), d, n, k, dp, dev(cen), dl)
    hip.synchronize()
    refined = ctx.labeling_refined_points()
    ctx.labeling_unpin_points()
    return dl.to_host()[:n], refined


@pytest.mark.parametrize("strategy", PRECISIONS)
@pytest.mark.parametrize("n,d,k,sigma,box", PIN_CASES)
def test_mfma_pinned_points_match_unpinned(gpu_ctx, strategy, n, d, k, sigma, box):
    # The copy holds exactly the operands and X2 the kernel would build from
    # the FP32 points, so the scores, the margin tests (the refined count)
    # and the labels are the same; the labels are the oracle's.  (The score
    # offset a call starts from is the previous call's largest: the
    # reference call follows one on the same data, as the pinned call does.)
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=n + d)
    cen = np.ascontiguousarray(pts[:, :k]) if n >= k else np.ascontiguousarray(pts[:, np.arange(k) % n])
    mfma_labels(gpu_ctx, strategy, pts, cen)
    ref = mfma_labels(gpu_ctx, strategy, pts, cen)
    ref_refined = gpu_ctx.labeling_refined_points()
    got, refined = pinned_labels(gpu_ctx, strategy, dev(pts), n, d, cen)
    np.testing.assert_array_equal(got, ref)
    assert refined == ref_refined
    sample = np.arange(0, n, 5)
    np.testing.assert_array_equal(got[sample], O.label(np.ascontiguousarray(pts[:, sample]), cen))


@pytest.mark.parametrize("strategy", PRECISIONS)
def test_mfma_pinned_special_values(gpu_ctx, strategy):
    pts, centres = clkmeans.generate_blobs(4096, 6, 12, sigma=1.0, box=10.0, seed=13)
    pts = pts.copy()
    pts[:, 100:164] *= np.float32(1e19)     # one step of huge points (exact scan)
    pts[:, 500] = np.nan
    pts[:, 501] = np.inf
    got, _ = pinned_labels(gpu_ctx, strategy, dev(pts), 4096, 6, centres)
    np.testing.assert_array_equal(got, O.label(pts, centres))


def test_mfma_pin_is_used_only_for_its_points(gpu_ctx):
    # Another buffer, another n, or the other 16-bit format: the labeling
    # reads the FP32 points, so the labels stay the oracle's.
    n, d, k = 5000, 24, 40
    pts, centres = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=5)
    other, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=6)
    dp, do = dev(pts), dev(other)
    want_o = O.label(other, centres)
    dl = zeros(n, np.uint32)
    gpu_ctx.labeling_pin_points("mfma_fp16", d, n, dp)
    for strategy, buf, nn, want in [("mfma_fp16", do, n, want_o),            # other buffer
                                    ("mfma_bf16", dp, n, O.label(pts, centres)),  # other format
                                    # the same buffer read as [d][n - 100]
                                    ("mfma_fp16", dp, n - 100, O.label(pts.ravel()[:d * (n - 100)].reshape(d, n - 100),
                                                                       centres))]:
        gpu_ctx.labeling(gpu_ctx.config(LAB, strategy), d, nn, k, buf, dev(centres), dl)
        hip.synchronize()
        np.testing.assert_array_equal(dl.to_host()[:nn], want)
    gpu_ctx.labeling_unpin_points()


def test_mfma_pin_rejects_bad_arguments(gpu_ctx):
    pts = dev(np.zeros((4, 100), np.float32))
    with pytest.raises(clkmeans.InvalidArgument):
        gpu_ctx.labeling_pin_points("unroll_vector", 4, 100, pts)
    with pytest.raises(clkmeans.ClkError):
        gpu_ctx.labeling_pin_points("mfma_fp16", 129, 100, pts)
    with pytest.raises(clkmeans.ClkError):
        gpu_ctx.labeling_pin_points("mfma_fp16", 4, 0, pts)
