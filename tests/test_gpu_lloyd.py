"""End-to-end Lloyd iterations through the C ABI against the oracle.

The workload of the CPU reference configuration: N=2^20 points, d=4, k=10,
sigma=1 blobs with centres in [-50,50]^4, seed 0, init = first 10 points,
30 iterations; single_stage (fused) and three_stage protocols.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import centroid_rel_err, dev, zeros

pytestmark = pytest.mark.gpu


def lloyd_gpu(ctx, pts, init, iterations, pipeline, check_each=False):
    d, n = pts.shape
    k = init.shape[1]
    dp = dev(pts)
    cen = [dev(init), zeros((d, k), np.float32)]
    sums = zeros((d, k), np.float64)
    masses = zeros(k, np.uint32)
    labels = zeros(n, np.uint32)
    fcfg = ctx.config(clkmeans.STAGE_FUSED, "cluster_merge")
    lcfg = ctx.config(clkmeans.STAGE_LABELING, "unroll_vector")
    mcfg = ctx.config(clkmeans.STAGE_MASS_UPDATE, "part_private")
    ccfg = ctx.config(clkmeans.STAGE_CENTROID_UPDATE, "cluster_merge")
    for it in range(iterations):
        old, new = cen[0], cen[1]
        if check_each:
            c_in = old.to_host()
        ctx.fill_zero(masses)
        ctx.fill_zero(sums)
        if pipeline == "single_stage_step":
            ctx.lloyd_step(fcfg, d, n, k, dp, old, new, sums, labels, masses)
        elif pipeline == "single_stage":
            ctx.fused(fcfg, d, n, k, dp, old, sums, labels, masses)
        else:
            ctx.labeling(lcfg, d, n, k, dp, old, labels)
            ctx.mass_update(mcfg, n, k, labels, masses)
            ctx.centroid_update(ccfg, d, n, k, dp, sums, labels, masses)
        if pipeline != "single_stage_step":
            ctx.divide(d, k, sums, masses, new)
        if check_each:
            hip.synchronize()
            lab = labels.to_host()
            ref = O.label(pts, c_in)
            assert np.array_equal(lab, ref), "iteration %d labels differ" % it
            m = masses.to_host()
            assert np.array_equal(m, O.mass(ref, k))
            ref_c = O.divide_from_f64(O.sums_f64(pts, ref, k), m)
            assert centroid_rel_err(new.to_host(), ref_c.astype(np.float64), pts, ref) <= 1e-5
        cen = [new, old]
    hip.synchronize()
    return cen[0].to_host(), labels.to_host(), masses.to_host()


@pytest.fixture(scope="module")
def c1_data():
    pts, _ = clkmeans.generate_blobs(1 << 20, 4, 10, sigma=1.0, box=50.0, seed=0)
    init = O.first_x(pts, 10)
    ref = O.lloyd(pts, init, 30)
    return pts, init, ref


@pytest.mark.parametrize("pipeline", ["single_stage", "single_stage_step", "three_stage"])
def test_lloyd_30_iterations_c1(gpu_ctx, c1_data, pipeline):
    pts, init, (rc, rl, rm) = c1_data
    c, l, m = lloyd_gpu(gpu_ctx, pts, init, 30, pipeline)
    np.testing.assert_array_equal(l, rl)
    np.testing.assert_array_equal(m, rm)
    ref64 = O.sums_f64(pts, rl, 10) / rm.astype(np.float64)
    assert centroid_rel_err(c, ref64, pts, rl) <= 1e-5


def test_lloyd_each_iteration_pinned(gpu_ctx):
    """Labels/masses bit-exact at every iteration given the GPU's own input
    centroids; a poor init (all seeds from few blobs) keeps margins small."""
    pts, _ = clkmeans.generate_blobs(1 << 16, 4, 10, sigma=3.0, box=10.0, seed=5)
    init = pts[:, ::4][:, :10].copy()
    lloyd_gpu(gpu_ctx, pts, init, 12, "single_stage", check_each=True)
    lloyd_gpu(gpu_ctx, pts, init, 12, "single_stage_step", check_each=True)
    lloyd_gpu(gpu_ctx, pts, init, 6, "three_stage", check_each=True)


def test_lloyd_matches_naive_kahan_labels(gpu_ctx, c1_data):
    """Against kmeans_naive's own arithmetic (Kahan FP32 sums): labels equal."""
    pts, init, _ = c1_data
    nc, nl, nm = O.lloyd(pts, init, 30, mode=O.MODE_NAIVE_KAHAN)
    c, l, m = lloyd_gpu(gpu_ctx, pts, init, 30, "single_stage")
    np.testing.assert_array_equal(l, nl)
    np.testing.assert_array_equal(m, nm)
