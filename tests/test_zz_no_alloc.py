"""clk_lloyd_step_noalloc (the entry point that allocates nothing) against
clk_lloyd_step: the same labels, masses, sums, centroids and shift, on two
calls in a row with the second input smaller, with the sum grid computed per
call (in the static scratch) and pinned on the context.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

FUSED = clkmeans.STAGE_FUSED


def iterate(ctx, noalloc, strategy, pts, init, iters=2):
    d, n = pts.shape
    k = init.shape[1]
    dp = dev(pts)
    cen = [dev(init), zeros((d, k), np.float32)]
    sums, labels, masses = zeros((d, k), np.float64), zeros(n, np.uint32), zeros(k, np.uint32)
    shift = zeros(1, np.float64)
    cfg = ctx.config(FUSED, strategy)
    step = ctx.lloyd_step_noalloc if noalloc else ctx.lloyd_step
    for _ in range(iters):
        step(cfg, d, n, k, dp, cen[0], cen[1], sums, labels, masses, max_shift=shift)
        cen.reverse()
    hip.synchronize()
    return [cen[0].to_host(), labels.to_host(), masses.to_host(), sums.to_host(), shift.to_host()]


CASES = [  # the second call of each pair is smaller than the first
    (1 << 20, 4, 10, "cluster_merge_bounded"),
    (50001, 3, 7, "cluster_merge"),
    (300000, 32, 256, "cluster_merge"),  # the shared-tile path
    (20000, 8, 40, "feature_sum"),
]


@pytest.mark.parametrize("pinned", [False, True])
def test_noalloc_equals_lloyd_step(gpu_ctx, pinned):
    for n, d, k, strategy in CASES:
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=2.0, box=10.0, seed=n + k)
        init = O.first_x(pts, k)
        if pinned:
            gpu_ctx.sum_grid_from_max(d, n, np.abs(pts).max(axis=1))
        try:
            ref = iterate(gpu_ctx, False, strategy, pts, init)
            got = iterate(gpu_ctx, True, strategy, pts, init)
        finally:
            gpu_ctx.sum_grid_clear()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        # and the reference iteration: labels of the last step's input
        # centroids, exact grid sums
        _, lab1, _ = O.lloyd(pts, init, 1, mode=O.MODE_GRID)
        c2, l2, m2 = O.lloyd(pts, init, 2, mode=O.MODE_GRID)
        np.testing.assert_array_equal(got[1], l2)
        np.testing.assert_array_equal(got[2], m2)
        np.testing.assert_array_equal(got[0], c2)


def test_noalloc_rejects_the_staged_path(gpu_ctx):
    n, d, k = 4096, 64, 3000  # k*d beyond the LDS tiles: the separate stages would allocate
    pts = np.random.default_rng(1).standard_normal((d, n)).astype(np.float32)
    init = np.ascontiguousarray(pts[:, :k])
    with pytest.raises(clkmeans.ClkError) as e:
        iterate(gpu_ctx, True, "cluster_merge", pts, init, iters=1)
    assert e.value.code == clkmeans.CLK_ERR_UNSUPPORTED
