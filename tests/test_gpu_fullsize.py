"""The configurations at their real sizes.

  - 2^27 points, d=4, k=10 (2 GiB, the headline workload): two Lloyd
    iterations from first-x with the exhaustive and the bounded fused pass;
    every label against the oracle's scan, masses exact, centroids bit-exact
    against the oracle's grid sums and within 1e-5 of FP64 sums.
  - Serving with a large job (N=2^24, d=32, k=256) among small ones: every
    served result bit-identical to the plain per-job loop (clk_serve_serial),
    the large job's labels on a 65,536-point sample against the oracle's
    scan, its masses and centroids against the oracle's grid sums.
"""
import os

import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import centroid_rel_err, dev, zeros

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.mark.parametrize("strategy", ["cluster_merge", "cluster_merge_bounded"])
def test_headline_size_two_iterations(gpu_ctx, strategy):
    n, d, k = 1 << 27, 4, 10
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=50.0, seed=0, threads=THREADS)
    init = O.first_x(pts, k)
    dp = dev(pts)
    cen = [dev(init), zeros((d, k), np.float32)]
    sums, labels, masses = zeros((d, k), np.float64), zeros(n, np.uint32), zeros(k, np.uint32)
    cfg = gpu_ctx.config(clkmeans.STAGE_FUSED, strategy)
    scales = O.grid_scales(pts)
    gpu_ctx.sum_grid_from_points(d, n, dp)
    try:
        for it in range(2):
            c_in = cen[0].to_host()
            gpu_ctx.lloyd_step(cfg, d, n, k, dp, cen[0], cen[1], sums, labels, masses)
            hip.synchronize()
            lab = labels.to_host()
            _, ref_lab, _ = O.lloyd_mt(pts, c_in, 1, THREADS)  # labels of c_in
            assert np.array_equal(lab, ref_lab), "iteration %d: labels differ" % it
            m = masses.to_host()
            np.testing.assert_array_equal(m, np.bincount(lab, minlength=k))
            ref = O.sums_grid(pts, lab, k, scales)
            np.testing.assert_array_equal(sums.to_host(), ref)
            got_c = cen[1].to_host()
            np.testing.assert_array_equal(got_c, O.divide_from_f64(ref, m))
            ref64 = O.sums_f64(pts, lab, k) / m.astype(np.float64)
            assert centroid_rel_err(got_c, ref64, pts, lab) <= 1e-5
            cen.reverse()
    finally:
        gpu_ctx.sum_grid_clear()


def test_serving_with_a_large_job(gpu_ctx):
    s = hip.Stream()
    specs = [x for x in clkmeans.job_mix(12, seed=41, log2n=(10, 16)) if x[1] <= 32][:8]
    big = (1 << 24, 32, 256)
    big_pts, _ = clkmeans.generate_blobs(*big, sigma=1.0, box=10.0, seed=99, threads=THREADS)
    big_init = O.first_x(big_pts, big[2])
    arrivals = clkmeans.poisson_arrivals(len(specs) + 1, 2000.0, seed=8)

    def requests():
        reqs = []
        for i, (n, d, k, seed) in enumerate(specs):
            pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=seed)
            reqs.append(clkmeans.ServeRequest(pts, O.first_x(pts, k), arrivals[i + (i >= 3)]))
        # the large job arrives fourth, while small ones are resident
        reqs.insert(3, clkmeans.ServeRequest(big_pts, big_init, arrivals[3]))
        return reqs

    served = requests()
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=5 << 30, max_d=32, max_k=256, labeling="mfma_bf16")
    stats = clkmeans.serve(b, served, s.handle)
    b.close()
    assert stats["max_active"] > 1
    serial = requests()
    clkmeans.serve_serial(gpu_ctx, serial, s.handle)
    for a, r in zip(served, serial):
        assert a.iterations == r.iterations and a.shift == r.shift
        np.testing.assert_array_equal(a.out_centroids.view(np.uint32), r.out_centroids.view(np.uint32))
        np.testing.assert_array_equal(a.out_labels, r.out_labels)
        np.testing.assert_array_equal(a.out_masses, r.out_masses)
    L = served[3]
    assert L.n == big[0]
    # the large job's last labeling used the centroids of iteration it-1:
    # the same job alone for it-1 iterations gives them (results are exact)
    one = clkmeans.Batch(gpu_ctx, max_jobs=1, arena_bytes=4 << 30, max_d=32, max_k=256, labeling="mfma_bf16")
    slot = one.add(big_pts, big_init, max_iterations=L.iterations - 1, stream=s.handle)
    for _ in range(L.iterations - 1):
        one.step(s.handle)
    c_prev = one.take(slot, s.handle)[0]
    one.close()
    idx = np.sort(np.random.default_rng(0).choice(big[0], 1 << 16, replace=False))
    np.testing.assert_array_equal(L.out_labels[idx], O.label(np.ascontiguousarray(big_pts[:, idx]), c_prev))
    np.testing.assert_array_equal(L.out_masses, np.bincount(L.out_labels, minlength=big[2]))
    ref = O.sums_grid(big_pts, L.out_labels, big[2])
    with np.errstate(invalid="ignore", divide="ignore"):
        np.testing.assert_array_equal(L.out_centroids, O.divide_from_f64(ref, L.out_masses))
    s.synchronize()
