"""Serving (clk_serve_run): jobs arrive over time, are admitted as slots and
arena space free up, iterate together with the resident jobs, and leave at
their tolerance or iteration limit.  Each result must be bit-identical to the
same job run alone in a batch, and the iterations must follow the reference
loop (labels and masses exact against the oracle's Lloyd loop).
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import centroid_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    s = hip.Stream()
    yield s
    s.synchronize()


def make_requests(specs, rate, seed, pinned=None, tol=1e-4, max_iter=50):
    arrivals = clkmeans.poisson_arrivals(len(specs), rate, seed)
    reqs = []
    for (n, d, k, s), t in zip(specs, arrivals):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=s)
        init = O.first_x(pts, k)
        if pinned is not None:  # page-locked copies (asynchronous admission)
            p = hip.PinnedArray(pts.shape, np.float32)
            p.array[...] = pts
            pinned.append(p)
            pts = p.array
        reqs.append(clkmeans.ServeRequest(pts, init, t, max_iterations=max_iter, tolerance=tol))
    return reqs


def run_alone(ctx, r, stream, labeling, max_d=32, max_k=256):
    b = clkmeans.Batch(ctx, max_jobs=1, arena_bytes=64 << 20, max_d=max_d, max_k=max_k, labeling=labeling)
    s = b.add(r.points, r.centroids, r.max_iterations, tolerance=r.tolerance, stream=stream.handle)
    for _ in range(r.max_iterations):
        b.step(stream.handle)
        if b.poll(stream.handle)[1] == 0:
            break
    out = b.take(s, stream.handle)
    b.close()
    return out


SPECS = [s for s in clkmeans.job_mix(40, seed=11, log2n=(10, 15)) if s[1] <= 32][:24]


@pytest.mark.parametrize("labeling", ["mfma_bf16", "unroll_vector"])
def test_served_results_equal_alone(gpu_ctx, stream, labeling):
    pinned = []
    # ~24 jobs within ~1 ms: more than the slots, so admission has to wait
    reqs = make_requests(SPECS, rate=20000.0, seed=5, pinned=pinned)
    # 6 slots: later jobs wait for earlier ones to leave and join mid-flight
    b = clkmeans.Batch(gpu_ctx, max_jobs=6, arena_bytes=128 << 20, max_d=32, max_k=256, labeling=labeling)
    stats = clkmeans.serve(b, reqs, stream.handle)
    b.close()
    assert stats["max_active"] > 1 and stats["steps"] > 0
    assert all(r.finished >= r.admitted >= r.arrival for r in reqs)
    assert any(r.admitted - r.arrival > 1e-3 for r in reqs)  # some waited for a slot
    for r in reqs:
        c, l, m, it, shift = run_alone(gpu_ctx, r, stream, labeling)
        assert r.iterations == it
        assert r.shift == shift
        np.testing.assert_array_equal(r.out_centroids.view(np.uint32), c.view(np.uint32))
        np.testing.assert_array_equal(r.out_labels, l)
        np.testing.assert_array_equal(r.out_masses, m)
        assert it == r.max_iterations or shift < r.tolerance
    for r in reqs[:6]:  # the reference iteration, exactly
        rc, rl, rm = O.lloyd(np.ascontiguousarray(r.points), r.centroids, r.iterations)
        np.testing.assert_array_equal(r.out_labels, rl)
        np.testing.assert_array_equal(r.out_masses, rm)
        with np.errstate(invalid="ignore", divide="ignore"):
            ref64 = O.sums_f64(np.ascontiguousarray(r.points), rl, r.k) / rm.astype(np.float64)
        assert centroid_rel_err(r.out_centroids, ref64, r.points, rl) <= 1e-5


def test_served_admission_by_arena(gpu_ctx, stream):
    # all jobs arrive at once; the arena holds about two at a time
    specs = [(20000, 16, 12, 70 + i) for i in range(5)]
    reqs = make_requests(specs, rate=1e6, seed=1, tol=0.0, max_iter=6)
    probe = clkmeans.Batch(gpu_ctx, 1, 1 << 20, 16, 16)
    one = probe.job_bytes(20000, 16, 12)
    probe.close()
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=int(one * 2.5), max_d=16, max_k=16)
    stats = clkmeans.serve(b, reqs, stream.handle)
    b.close()
    assert stats["max_active"] == 2
    order = sorted(reqs, key=lambda r: r.admitted)
    assert order[2].admitted >= min(order[0].finished, order[1].finished)  # waited for space
    for r in reqs:
        assert r.iterations == 6
        rc, rl, rm = O.lloyd(r.points, r.centroids, 6)
        np.testing.assert_array_equal(r.out_labels, rl)
        np.testing.assert_array_equal(r.out_masses, rm)


def test_serial_loop_follows_the_reference(gpu_ctx, stream):
    reqs = make_requests(SPECS[:6], rate=1000.0, seed=2)
    stats = clkmeans.serve_serial(gpu_ctx, reqs, stream.handle)
    assert stats["max_active"] == 1 and stats["solves_per_s"] > 0
    prev = 0.0
    for r in reqs:
        assert r.admitted >= max(r.arrival, prev) - 1e-9  # one job at a time, in order
        prev = r.finished
        assert r.iterations == r.max_iterations or r.shift < r.tolerance
        rc, rl, rm = O.lloyd(r.points, r.centroids, r.iterations)
        np.testing.assert_array_equal(r.out_labels, rl)
        np.testing.assert_array_equal(r.out_masses, rm)


def test_serve_rejects_bad_input(gpu_ctx, stream):
    reqs = make_requests(SPECS[:2], rate=100.0, seed=3)
    reqs[0].arrival, reqs[1].arrival = 1.0, 0.5  # not in arrival order
    b = clkmeans.Batch(gpu_ctx, max_jobs=2, arena_bytes=64 << 20, max_d=32, max_k=256)
    with pytest.raises(clkmeans.InvalidArgument):
        clkmeans.serve(b, reqs, stream.handle)
    small = clkmeans.Batch(gpu_ctx, max_jobs=2, arena_bytes=1 << 16, max_d=32, max_k=256)
    reqs = make_requests(SPECS[:1], rate=100.0, seed=3)
    with pytest.raises(clkmeans.ClkError) as e:
        clkmeans.serve(small, reqs, stream.handle)  # larger than the whole arena: never admissible
    assert e.value.code == clkmeans.CLK_ERR_NOMEM
    small.close()
    b.close()


@pytest.mark.parametrize("labeling", ["mfma_bf16", "unroll_vector"])
def test_lone_job_path_equals_batch_path(gpu_ctx, stream, labeling):
    """At low load a lone job iterates through the per-job fused pass and
    switches back to the batch step when others arrive (MFMA jobs get their
    16-bit table rebuilt from their current centroids).  Whatever path ran
    each iteration, every result equals the same job alone in a batch."""
    # MFMA-class jobs (k >= 16, k*d >= 2048) and FP32-class jobs; long runs
    # (no tolerance) so that later arrivals meet jobs that iterated alone
    specs = [(20000, 32, 100, 3), (6000, 3, 12, 4), (30000, 16, 200, 5), (9000, 8, 40, 6), (4096, 32, 64, 7),
             (12000, 8, 300, 8)]
    reqs = []
    # arrivals spread over ~5 ms, each job taking ~1-2 ms: the first runs
    # alone for a while, later ones overlap, the last finishes alone
    for (n, d, k, s), t in zip(specs, [0.0, 0.0004, 0.0007, 0.0025, 0.0027, 0.005]):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=2.0, box=10.0, seed=s)
        reqs.append(clkmeans.ServeRequest(pts, O.first_x(pts, k), t, max_iterations=40, tolerance=0.0))
    b = clkmeans.Batch(gpu_ctx, max_jobs=4, arena_bytes=256 << 20, max_d=32, max_k=300, labeling=labeling)
    stats = clkmeans.serve(b, reqs, stream.handle)
    b.close()
    assert stats["solo_steps"] > 0 and stats["steps"] > stats["solo_steps"]
    for r in reqs:
        c, l, m, it, shift = run_alone(gpu_ctx, r, stream, labeling, max_k=300)
        assert r.iterations == it == 40
        assert r.shift == shift
        np.testing.assert_array_equal(r.out_centroids.view(np.uint32), c.view(np.uint32))
        np.testing.assert_array_equal(r.out_labels, l)
        np.testing.assert_array_equal(r.out_masses, m)
    # and the plain per-job loop
    sreqs = [clkmeans.ServeRequest(r.points, r.centroids, r.arrival, max_iterations=40, tolerance=0.0)
             for r in reqs]
    clkmeans.serve_serial(gpu_ctx, sreqs, stream.handle)
    for a, r in zip(reqs, sreqs):
        np.testing.assert_array_equal(a.out_centroids.view(np.uint32), r.out_centroids.view(np.uint32))
        np.testing.assert_array_equal(a.out_labels, r.out_labels)
