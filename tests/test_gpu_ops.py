"""Parity of every C-ABI operator against the CPU oracle, on the GPU.

Bar: labels and masses bit-exact; FP64 centroid sums within a bound derived
from FP32 per-work-item accumulation; divide bit-exact given equal inputs.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import centroid_rel_err, dev, zeros

pytestmark = pytest.mark.gpu

FUSED = clkmeans.STAGE_FUSED
LAB = clkmeans.STAGE_LABELING
MASS = clkmeans.STAGE_MASS_UPDATE
CU = clkmeans.STAGE_CENTROID_UPDATE


def blobs(n, d, k, seed=0, sigma=1.0, box=50.0):
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=seed)
    return pts


def sum_tolerance(points, labels, k, per_item=None):
    """|gpu - fp64| allowed per (feature, cluster) cell: 1e-5 of the cell's
    sum of |x| -- the 1e-5 relative centroid bar, measured against the
    cluster's mean |x| so that coordinates near 0 do not inflate it."""
    d = points.shape[0]
    absx = np.zeros((d, k))
    for f in range(d):
        absx[f] = np.bincount(labels, weights=np.abs(points[f].astype(np.float64)), minlength=k)
    return 1e-5 * absx + 1e-30


def run_fused(ctx, strategy, pts, cen, **kw):
    d, n = pts.shape
    k = cen.shape[1]
    dp, dc = dev(pts), dev(cen)
    sums0 = np.arange(d * k, dtype=np.float64).reshape(d, k) * 0.5  # accumulate semantics
    mass0 = np.arange(k, dtype=np.uint32) * 3
    ds, dm = dev(sums0), dev(mass0)
    dl = zeros(max(n, 1), np.uint32)
    cfg = ctx.config(FUSED, strategy, **kw)
    ctx.fused(cfg, d, n, k, dp, dc, ds, dl, dm)
    hip.synchronize()
    return dl.to_host()[:n], ds.to_host() - sums0, dm.to_host() - mass0


FUSED_CASES = [
    # n, d, k, kwargs
    (1 << 16, 4, 10, {}),
    ((1 << 16) + 3, 4, 10, {}),          # n % 4 != 0 -> scalar loads
    (1000, 4, 10, {"vector_length": 2}),
    (777, 3, 5, {}),                      # non-power-of-two d (padded bucket)
    (4096, 1, 4, {}),
    (4096, 2, 1, {}),                     # k == 1
    (5000, 8, 37, {}),
    (3000, 16, 20, {"local_size": 128}),
    (2000, 5, 7, {"local_size": 64, "vector_length": 1}),
    (2048, 64, 8, {}),
    (1500, 100, 6, {}),                   # d > 64: global path
    (4000, 64, 700, {}),                  # k*d beyond LDS: labeling + mass + slice stages
    (20000, 4, 10, {"global_size": 4096, "local_size": 32, "vector_length": 4}),  # GPU .conf geometry
    (7, 4, 10, {}),                       # n < k
]


@pytest.mark.parametrize("strategy", ["cluster_merge", "feature_sum"])
@pytest.mark.parametrize("n,d,k,kw", FUSED_CASES)
def test_fused_matches_oracle(gpu_ctx, strategy, n, d, k, kw):
    pts = blobs(n, d, k, seed=n + d + k)
    cen = O.first_x(pts, k) if n >= k else np.random.default_rng(1).standard_normal((d, k)).astype(np.float32)
    labels, sums, masses = run_fused(gpu_ctx, strategy, pts, cen, **kw)
    ref_lab = O.label(pts, cen)
    np.testing.assert_array_equal(labels, ref_lab)
    np.testing.assert_array_equal(masses, O.mass(ref_lab, k))
    ref_sums = O.sums_f64(pts, ref_lab, k)
    tol = sum_tolerance(pts, ref_lab, k, per_item=max(1, n))
    assert np.all(np.abs(sums - ref_sums) <= tol)


def test_fused_nan_and_ties(gpu_ctx):
    """NaN centroids (empty clusters) never win; equal distances -> lowest index."""
    rng = np.random.default_rng(7)
    pts = rng.integers(-3, 4, size=(4, 4096)).astype(np.float32)
    cen = rng.integers(-3, 4, size=(4, 12)).astype(np.float32)
    cen[:, 5] = cen[:, 2]          # exact duplicate -> ties resolved to 2
    cen[:, 0] = np.nan             # empty cluster from a previous iteration
    cen[1, 7] = np.nan
    for strategy in ("cluster_merge", "feature_sum"):
        labels, sums, masses = run_fused(gpu_ctx, strategy, pts, cen)
        ref = O.label(pts, cen)
        np.testing.assert_array_equal(labels, ref)
        assert not np.any(labels == 5) and not np.any(labels == 0) and not np.any(labels == 7)
        np.testing.assert_array_equal(masses, O.mass(ref, 12))


def test_fused_all_nan_gives_label_zero(gpu_ctx):
    pts = blobs(1024, 4, 3)
    cen = np.full((4, 3), np.nan, np.float32)
    labels, _, masses = run_fused(gpu_ctx, "cluster_merge", pts, cen)
    assert np.all(labels == 0) and masses[0] == 1024


def test_fused_zero_points_is_noop(gpu_ctx):
    pts = np.zeros((4, 0), np.float32)
    cen = np.zeros((4, 3), np.float32)
    labels, sums, masses = run_fused(gpu_ctx, "cluster_merge", pts, cen)
    assert labels.size == 0 and sums.sum() == 0 and masses.sum() == 0


@pytest.mark.parametrize("n,d,k,vec", [(1 << 16, 4, 10, 4), (999, 3, 17, 4), (4096, 32, 50, 2), (800, 80, 5, 1),
                                       (2048, 64, 2000, 4)])
def test_labeling_matches_oracle(gpu_ctx, n, d, k, vec):
    pts = blobs(n, d, min(k, 64), seed=3)
    cen = np.random.default_rng(k).standard_normal((d, k)).astype(np.float32) * 20
    dl = zeros(n, np.uint32)
    gpu_ctx.labeling(gpu_ctx.config(LAB, "unroll_vector", vector_length=vec), d, n, k, dev(pts), dev(cen), dl)
    hip.synchronize()
    np.testing.assert_array_equal(dl.to_host(), O.label(pts, cen))


@pytest.mark.parametrize("n,d,k,vec", [(4096, 64, 1024, 4), (3001, 30, 2000, 4), (2048, 8, 6000, 2),
                                       (1000, 3, 20000, 4)])
def test_labeling_centroid_chunks_match_oracle(gpu_ctx, n, d, k, vec):
    """k*d beyond LDS: centroids pass through LDS in chunks (64 KiB); a tie
    across a chunk boundary keeps the lower index, NaN centroids never win."""
    pts = blobs(n, d, 32, seed=5)
    rng = np.random.default_rng(d + k)
    cen = (rng.standard_normal((d, k)) * 5).astype(np.float32)
    chunk = (64 << 10) // (4 * {3: 4, 8: 8, 30: 32, 64: 64}[d])
    cen[:, chunk + 3] = cen[:, 3]            # duplicate in the next chunk: label 3 wins
    cen[:, 1] = pts[:, 0]                    # point 0 sits on centroid 1 (and on its copy below)
    cen[:, chunk + 1] = pts[:, 0]
    cen[:, chunk - 1] = np.nan               # an empty cluster at a chunk's end
    dl = zeros(n, np.uint32)
    gpu_ctx.labeling(gpu_ctx.config(LAB, "unroll_vector", vector_length=vec), d, n, k, dev(pts), dev(cen), dl)
    hip.synchronize()
    got = dl.to_host()
    np.testing.assert_array_equal(got, O.label(pts, cen))
    assert got[0] == 1 and not np.any(got == chunk - 1) and not np.any(got == chunk + 3)


@pytest.mark.parametrize("strategy", ["global_atomic", "part_global", "part_local", "part_private"])
@pytest.mark.parametrize("n,k,kw", [(1 << 18, 10, {}), (100003, 7, {}), (5000, 1000, {}),
                                    (64, 3, {"global_size": 128, "local_size": 1, "vector_length": 8}),
                                    (4096, 4, {"global_size": 32768, "local_size": 32, "vector_length": 4})])
def test_mass_update_exact(gpu_ctx, strategy, n, k, kw):
    rng = np.random.default_rng(n + k)
    labels = rng.integers(0, k, size=n).astype(np.uint32)
    m0 = rng.integers(0, 1000, size=k).astype(np.uint32)
    dm = dev(m0)
    gpu_ctx.mass_update(gpu_ctx.config(MASS, strategy, **kw), n, k, dev(labels), dm)
    hip.synchronize()
    np.testing.assert_array_equal(dm.to_host(), m0 + O.mass(labels, k))


@pytest.mark.parametrize("strategy", ["feature_sum", "feature_sum_pardim", "cluster_merge"])
@pytest.mark.parametrize("n,d,k,kw", [(1 << 16, 4, 10, {}), (9999, 3, 13, {}),
                                      (4096, 4, 10, {"local_size": 64, "local_features": 4,
                                                     "thread_features": 1, "vector_length": 4}),
                                      (3000, 16, 30, {"local_features": 2, "thread_features": 3}),
                                      (2000, 64, 300, {}), (1000, 70, 4, {}),
                                      (50000, 64, 1024, {}),   # k*d past LDS: feature slices
                                      (30001, 45, 1000, {})])  # partial last slice
def test_centroid_update_matches_oracle(gpu_ctx, strategy, n, d, k, kw):
    pts = blobs(n, d, k, seed=11)
    labels = np.random.default_rng(5).integers(0, k, size=n).astype(np.uint32)
    s0 = np.full((d, k), 0.25)
    ds = dev(s0)
    gpu_ctx.centroid_update(gpu_ctx.config(CU, strategy, **kw), d, n, k, dev(pts), ds, dev(labels))
    hip.synchronize()
    got = ds.to_host() - s0
    ref = O.sums_f64(pts, labels, k)
    assert np.all(np.abs(got - ref) <= sum_tolerance(pts, labels, k, per_item=n))


@pytest.mark.parametrize("hot", [1, 3, 64])
def test_centroid_slices_repeated_labels(gpu_ctx, hot):
    # Large k*d (feature-slice path) with labels drawn from only `hot`
    # clusters: every 64-point wave carries many equal labels, so the owner
    # claim of each slice table resolves them over several rounds.
    n, d, k = 40000, 33, 1500
    pts = blobs(n, d, 8, seed=hot)
    labels = (np.random.default_rng(hot).integers(0, hot, size=n) * 7 % k).astype(np.uint32)
    ds = zeros((d, k), np.float64)
    gpu_ctx.centroid_update(gpu_ctx.config(CU, "cluster_merge"), d, n, k, dev(pts), ds, dev(labels))
    hip.synchronize()
    ref = O.sums_f64(pts, labels, k)
    assert np.all(np.abs(ds.to_host() - ref) <= sum_tolerance(pts, labels, k, per_item=n))


@pytest.mark.parametrize("n,d,k", [
    (1_500_000, 24, 2000),  # ~34 tiles per wave: full rounds of the 8-tile ring, then a partial one
    (10_000, 33, 1500),     # fewer tiles than waves: some waves get none
])
def test_centroid_slices_prefetch_ring(gpu_ctx, n, d, k):
    # The feature-slice kernel keeps 8 tiles per wave in flight; the loads
    # past a wave's last tile are clamped and must never be added.
    pts = blobs(n, d, 16, seed=d)
    labels = np.random.default_rng(d).integers(0, k, size=n).astype(np.uint32)
    s0 = np.full((d, k), -1.5)
    ds = dev(s0)
    gpu_ctx.centroid_update(gpu_ctx.config(CU, "cluster_merge"), d, n, k, dev(pts), ds, dev(labels))
    hip.synchronize()
    ref = O.sums_f64(pts, labels, k)
    assert np.all(np.abs(ds.to_host() - s0 - ref) <= sum_tolerance(pts, labels, k, per_item=n))


def test_divide_bit_exact_and_nan(gpu_ctx):
    rng = np.random.default_rng(2)
    d, k = 4, 10
    sums = rng.standard_normal((d, k)) * 1e4
    masses = rng.integers(1, 5000, size=k).astype(np.uint32)
    masses[3] = 0
    sums[:, 3] = 0.0
    old = rng.standard_normal((d, k)).astype(np.float32)
    dc = zeros((d, k), np.float32)
    shift = zeros(1, np.float64)
    gpu_ctx.divide(d, k, dev(sums), dev(masses), dc, old_centroids=dev(old), max_shift=shift)
    hip.synchronize()
    got = dc.to_host()
    ref = O.divide_from_f64(sums, masses)
    assert np.all(np.isnan(got[:, 3])) and np.all(np.isnan(ref[:, 3]))
    live = masses != 0
    np.testing.assert_array_equal(got[:, live].view(np.uint32), ref[:, live].view(np.uint32))
    ok = masses != 0
    ref_shift = np.sqrt(((ref[:, ok].astype(np.float64) - old[:, ok]) ** 2).sum(0)).max()
    assert abs(shift.to_host()[0] - ref_shift) <= 1e-12 * ref_shift


# The reference's ReduceVectorParcol tests (test/reduce_vector.cpp:74-221):
# shapes rows x cols, uint32 wrap-around, compared with a sequential sum.
REDUCE_CASES = {
    "Ascending": (4, 256, "ascending"), "Large": (4, 256, "large"), "Random": (4, 256, "random"),
    "BigRandom": (4, 2048 * 32, "random"), "ManyRows": (2048, 2048 * 32, "random"),
    "TeslaSize": (256, 90 * 32 * 16, "random"), "FurySize": (4, 56 * 64 * 16, "random"),
    "OddCols": (10, 769, "random"),
}


@pytest.mark.parametrize("case", sorted(REDUCE_CASES))
def test_reduce_vector_parcol_u32(gpu_ctx, case):
    rows, cols, kind = REDUCE_CASES[case]
    if kind == "ascending":
        data = np.arange(rows * cols, dtype=np.uint64).astype(np.uint32)
    elif kind == "large":
        data = np.full(rows * cols, 1 << 31, np.uint32)
    else:
        data = np.random.default_rng(rows * cols).integers(0, 2 ** 32, size=rows * cols, dtype=np.uint64).astype(np.uint32)
    dd = dev(data)
    gpu_ctx.reduce_vector_parcol(cols, rows, dd)
    hip.synchronize()
    np.testing.assert_array_equal(dd.to_host()[:rows], O.reduce_u32(data, cols, rows))


def test_reduce_vector_parcol_f64(gpu_ctx):
    rows, cols = 40, 768
    data = np.random.default_rng(0).standard_normal(rows * cols) * 100
    dd = dev(data)
    gpu_ctx.reduce_vector_parcol(cols, rows, dd)
    hip.synchronize()
    got = dd.to_host()[:rows]
    ref = np.array([np.sum(data.reshape(cols, rows)[:, r], dtype=np.longdouble) for r in range(rows)], np.float64)
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(data).reshape(cols, rows).sum(0))


def test_fill_first_x_and_add(gpu_ctx):
    buf = dev(np.arange(1001, dtype=np.uint8))
    gpu_ctx.fill_zero(buf)
    hip.synchronize()
    assert not buf.to_host().any()
    pts = blobs(100, 3, 5)
    dc = zeros((3, 7), np.float32)
    gpu_ctx.first_x(3, 100, 7, dev(pts), dc)
    hip.synchronize()
    np.testing.assert_array_equal(dc.to_host(), O.first_x(pts, 7))
    a = np.arange(50, dtype=np.uint32)
    da = dev(a)
    gpu_ctx.matrix_add(50, da, dev(a * 2))
    hip.synchronize()
    np.testing.assert_array_equal(da.to_host(), a * 3)


def test_errors_are_raised(gpu_ctx):
    with pytest.raises(ValueError):
        gpu_ctx.config(FUSED, "no_such_strategy")
    with pytest.raises(AssertionError):
        gpu_ctx.fused(gpu_ctx.config(FUSED, "cluster_merge"), 0, 10, 3, zeros(1, np.float32),
                      zeros(1, np.float32), zeros(1, np.float64), zeros(1, np.uint32), zeros(1, np.uint32))


@pytest.mark.parametrize("strategy", ["cluster_merge", "feature_sum"])
def test_lloyd_step_and_repeated_launches(gpu_ctx, strategy):
    """One-launch iteration (in-kernel merge + divide) for grids that are not
    multiples of the merge group, repeated so that the merge counters must
    have been reset by every launch; accumulate mode doubles on a 2nd call."""
    n, d, k = 100000, 4, 10
    pts = blobs(n, d, k, seed=21)
    cen = O.first_x(pts, k)
    ref_lab = O.label(pts, cen)
    ref_m = O.mass(ref_lab, k)
    ref_s = O.sums_f64(pts, ref_lab, k)
    tol = sum_tolerance(pts, ref_lab, k)
    dp, dc = dev(pts), dev(cen)
    for grid_blocks in (1, 15, 16, 17, 33, 400, 0, 17, 1):
        kw = {"global_size": grid_blocks * 256} if grid_blocks else {}
        cfg = gpu_ctx.config(FUSED, strategy, **kw)
        newc, sums, lab, mass = zeros((d, k), np.float32), dev(np.full((d, k), 7.0)), zeros(n, np.uint32), \
            dev(np.full(k, 5, np.uint32))
        shift = zeros(1, np.float64)
        gpu_ctx.lloyd_step(cfg, d, n, k, dp, dc, newc, sums, lab, mass, max_shift=shift)
        hip.synchronize()
        np.testing.assert_array_equal(lab.to_host(), ref_lab)
        m = mass.to_host()
        np.testing.assert_array_equal(m, ref_m)                    # overwritten, not accumulated
        s = sums.to_host()
        assert np.all(np.abs(s - ref_s) <= tol)
        np.testing.assert_array_equal(newc.to_host(), O.divide_from_f64(s, m))  # divide of its own sums
        ref_shift = np.sqrt(((O.divide_from_f64(s, m).astype(np.float64) - cen) ** 2).sum(0)).max()
        assert abs(shift.to_host()[0] - ref_shift) <= 1e-12 * ref_shift
        # accumulate mode, twice
        s2, m2 = zeros((d, k), np.float64), zeros(k, np.uint32)
        gpu_ctx.fused(cfg, d, n, k, dp, dc, s2, lab, m2)
        gpu_ctx.fused(cfg, d, n, k, dp, dc, s2, lab, m2)
        hip.synchronize()
        np.testing.assert_array_equal(m2.to_host(), 2 * ref_m)
        assert np.all(np.abs(s2.to_host() - 2 * ref_s) <= 2 * tol)


def test_fused_is_deterministic(gpu_ctx):
    """Fixed reduction order: identical FP64 sums bit for bit across runs."""
    n, d, k = 1 << 18, 4, 10
    pts = blobs(n, d, k, seed=3)
    cen = O.first_x(pts, k)
    dp, dc = dev(pts), dev(cen)
    outs = []
    for _ in range(3):
        newc, sums, lab, mass = zeros((d, k), np.float32), zeros((d, k), np.float64), zeros(n, np.uint32), \
            zeros(k, np.uint32)
        gpu_ctx.lloyd_step(gpu_ctx.config(FUSED, "cluster_merge"), d, n, k, dp, dc, newc, sums, lab, mass)
        hip.synchronize()
        outs.append(sums.to_host().view(np.uint64).copy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_lloyd_step_large_dk_uses_column_merge(gpu_ctx):
    """d*k past the one-work-group merge but within the shared LDS tile
    (d=32, k=600: 156 KiB tile, 150 KiB of FP64 totals)."""
    n, d, k = 20000, 32, 600
    pts = blobs(n, d, 64, seed=8)
    cen = np.random.default_rng(3).standard_normal((d, k)).astype(np.float32) * 20
    newc, sums, lab, mass = zeros((d, k), np.float32), zeros((d, k), np.float64), zeros(n, np.uint32), \
        zeros(k, np.uint32)
    gpu_ctx.lloyd_step(gpu_ctx.config(FUSED, "feature_sum"), d, n, k, dev(pts), dev(cen), newc, sums, lab, mass)
    hip.synchronize()
    ref = O.label(pts, cen)
    np.testing.assert_array_equal(lab.to_host(), ref)
    m = mass.to_host()
    np.testing.assert_array_equal(m, O.mass(ref, k))
    s = sums.to_host()
    assert np.all(np.abs(s - O.sums_f64(pts, ref, k)) <= sum_tolerance(pts, ref, k))
    got = newc.to_host()
    exp = O.divide_from_f64(s, m)
    live = m != 0
    np.testing.assert_array_equal(got[:, live], exp[:, live])
    assert np.all(np.isnan(got[:, ~live]))
