"""The batch engine (clkmeans_batch.h): many jobs per persistent launch.

Each job must follow the reference iteration exactly (labels and masses
bit-exact vs the oracle's FP64-merge Lloyd loop on the same inputs, centroids
within 1e-5 relative of FP64), its result must not depend on the other jobs
of the batch, and admission must respect the arena.
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import centroid_rel_err

pytestmark = pytest.mark.gpu


def make_job(n, d, k, seed, sigma=1.0, box=10.0):
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=seed)
    return pts, O.first_x(pts, k)


def check_job(pts, init, iters, got):
    c, l, m, it, _ = got
    k = init.shape[1]
    rc, rl, rm = O.lloyd(pts, init, iters)
    assert it == iters
    np.testing.assert_array_equal(l, rl)
    np.testing.assert_array_equal(m, rm)
    with np.errstate(invalid="ignore", divide="ignore"):
        ref64 = O.sums_f64(pts, rl, k) / rm.astype(np.float64)
    assert centroid_rel_err(c, ref64, pts, rl) <= 1e-5
    np.testing.assert_array_equal(np.isnan(c).all(axis=0), rm == 0)


@pytest.fixture(scope="module")
def stream():
    s = hip.Stream()
    yield s
    s.synchronize()


MIX = [  # n, d, k
    (3000, 3, 4), (1025, 8, 64), (20000, 16, 30), (7777, 32, 100), (50000, 3, 10),
    (4096, 8, 256), (12345, 16, 7), (1500, 32, 93), (65536, 8, 40), (999, 2, 5),
    (30000, 32, 256), (2048, 64, 12), (5000, 1, 3), (100, 5, 1),
]


LABELINGS = ["unroll_vector", "mfma_fp16", "mfma_bf16"]


@pytest.mark.parametrize("labeling", LABELINGS)
def test_batch_matches_oracle(gpu_ctx, stream, labeling):
    b = clkmeans.Batch(gpu_ctx, max_jobs=32, arena_bytes=256 << 20, max_d=64, max_k=256, labeling=labeling)
    jobs = []
    for i, (n, d, k) in enumerate(MIX):
        pts, init = make_job(n, d, k, seed=100 + i)
        jobs.append((b.add(pts, init, max_iterations=8, stream=stream.handle), pts, init))
    for _ in range(8):
        b.step(stream.handle)
    states, active = b.poll(stream.handle)
    assert active == 0 and all(states[s] == clkmeans.JOB_DONE for s, _, _ in jobs)
    for s, pts, init in jobs:
        check_job(pts, init, 8, b.take(s, stream.handle))
    b.close()


ORDER_JOBS = [  # n, d, k
    (20000, 3, 4), (15000, 16, 40), (9000, 32, 120), (30000, 8, 200), (4096, 32, 16), (7000, 2, 70),
]


@pytest.mark.parametrize("labeling", ["unroll_vector", "mfma_bf16"])
@pytest.mark.parametrize("order", ["interleaved", "grouped", "shuffled"])
def test_batch_point_orders(gpu_ctx, stream, labeling, order):
    # The tiles add a wave's grid integers lane by lane when its labels are
    # spread and along sorted runs when they repeat: points interleaved over
    # the blobs (k = 4: 16 lanes per label, the most same-cell atomics the
    # spread form takes), grouped by blob (runs of one label) and shuffled.
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=128 << 20, max_d=32, max_k=256, labeling=labeling)
    jobs = []
    for i, (n, d, k) in enumerate(ORDER_JOBS):
        pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=10.0, seed=300 + i)
        init = np.ascontiguousarray(pts[:, :k])  # one point of each blob
        if order == "grouped":
            pts = np.ascontiguousarray(pts[:, np.argsort(np.arange(n) % k, kind="stable")])
        elif order == "shuffled":
            pts = np.ascontiguousarray(pts[:, np.random.default_rng(i).permutation(n)])
        jobs.append((b.add(pts, init, max_iterations=5, stream=stream.handle), pts, init))
    for _ in range(5):
        b.step(stream.handle)
    for s, pts, init in jobs:
        check_job(pts, init, 5, b.take(s, stream.handle))
    b.close()


@pytest.mark.parametrize("labeling", ["unroll_vector", "mfma_bf16"])
def test_batch_keep_radius_paths(gpu_ctx, stream, labeling):
    # From the second iteration on: points inside their centroid's keep
    # radius keep their labels; the FP32 classes rescan a wave's few other
    # points with all lanes (many: per lane), the MFMA classes rescan one or
    # two exactly from the centroid rows (more: the MFMA path).  Jobs with
    # separated blobs (mostly kept), overlapping blobs (many rescans), a
    # duplicated initial centroid (no keep radius for either copy) and an
    # initial centroid far from every point (an empty cluster: NaN centroid);
    # with mfma_bf16 the k*d >= 2048 jobs take the MFMA classes (k <= 64
    # and k > 64 each).
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=256 << 20, max_d=32, max_k=256, labeling=labeling)
    jobs = []
    specs = [(40000, 4, 10, 1.0, 50.0), (40000, 3, 40, 2.0, 10.0), (30000, 8, 60, 1.0, 20.0),
             (20000, 16, 12, 3.0, 10.0), (25000, 32, 20, 1.0, 20.0), (30000, 16, 160, 1.0, 20.0),
             (30000, 32, 64, 2.0, 10.0)]
    for i, (n, d, k, sigma, box) in enumerate(specs):
        pts, init = make_job(n, d, k, seed=300 + i, sigma=sigma, box=box)
        if i in (2, 5):
            init[:, 5] = init[:, 3]           # duplicated centroid
            init[:, 7] = 1e6                  # far away: empty from the first iteration
        jobs.append((b.add(pts, init, max_iterations=12, stream=stream.handle), pts, init))
    for _ in range(12):
        b.step(stream.handle)
    for s, pts, init in jobs:
        check_job(pts, init, 12, b.take(s, stream.handle))
    b.close()


def test_batch_second_distance_bound(gpu_ctx, stream):
    # d <= 8 FP32 tiles also keep a point's label while its distance to its
    # centroid stays below its lower bound on every other centroid's distance
    # (shrunk each iteration by the largest centroid move).  Long runs on
    # overlapping blobs (the bound decides most points), a small k without
    # keep radii (the bound alone), lattice ties, a duplicated and an empty
    # (NaN) centroid, NaN points and points whose distances overflow.
    rng = np.random.default_rng(21)
    jobs = []
    for i, (n, d, k, sigma, box, iters) in enumerate([
            (60000, 3, 256, 1.0, 10.0, 25), (30000, 3, 12, 1.5, 5.0, 25), (40000, 8, 60, 2.0, 10.0, 20),
            (20000, 4, 30, 1.0, 10.0, 15), (25000, 6, 90, 1.0, 8.0, 20)]):
        pts, init = make_job(n, d, k, seed=500 + i, sigma=sigma, box=box)
        if i == 3:
            init[:, 5] = init[:, 3]  # duplicated centroid
            init[:, 7] = 1e6         # far away: empty, then NaN
        jobs.append((pts, init, iters))
    p = np.arange(-2, 9, dtype=np.float32)
    lat = np.stack(np.meshgrid(p, p, p, indexing="ij"), 0).reshape(3, -1)
    lat = np.ascontiguousarray(np.repeat(lat, 3, axis=1))
    g = np.arange(0, 7, 2, dtype=np.float32)
    cen = np.ascontiguousarray(np.stack(np.meshgrid(g, g, g, indexing="ij"), 0).reshape(3, -1))
    jobs.append((lat, cen, 10))  # every odd lattice coordinate ties two centroids
    pts, init = make_job(12000, 5, 20, seed=520)
    pts = pts.copy()
    pts[:, 41] = np.nan
    pts[:, 900:964] *= np.float32(1e19)
    pts[:, 2000] = np.inf
    jobs.append((pts, np.ascontiguousarray(pts[:, rng.choice(np.arange(3000, 12000), 20, replace=False)]), 12))
    res = []
    for serial in (False, True):  # the batch, and the plain loop on its buffers
        b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=256 << 20, max_d=8, max_k=256)
        try:
            slots = [b.add(p_, c_, it, stream=stream.handle) for p_, c_, it in jobs]
            if serial:
                b.run_serial(stream=stream.handle)
            else:
                for _ in range(25):
                    b.step(stream.handle)
            res.append([b.take(s, stream.handle) for s in slots])
        finally:
            b.close()
    for i, ((p_, c_, it), got, ser) in enumerate(zip(jobs, *res)):
        assert got[3] == it
        for x, y in zip(got[:4], ser[:4]):
            np.testing.assert_array_equal(x, y)
        if i < len(jobs) - 1:  # the oracle's grid loop (the last job holds NaN / inf)
            rc, rl, rm = O.lloyd(p_, c_, it, mode=O.MODE_GRID)
            np.testing.assert_array_equal(got[1], rl)
            np.testing.assert_array_equal(got[2], rm)
            np.testing.assert_array_equal(got[0], rc)


@pytest.mark.parametrize("labeling", LABELINGS)
def test_batch_result_independent_of_mix(gpu_ctx, stream, labeling):
    # the same job alone and among 20 others: identical bits
    pts, init = make_job(40000, 16, 37, seed=7)
    alone = clkmeans.Batch(gpu_ctx, max_jobs=4, arena_bytes=64 << 20, max_d=32, max_k=256, labeling=labeling)
    s = alone.add(pts, init, max_iterations=6, stream=stream.handle)
    for _ in range(6):
        alone.step(stream.handle)
    r_alone = alone.take(s, stream.handle)
    alone.close()
    mixed = clkmeans.Batch(gpu_ctx, max_jobs=32, arena_bytes=256 << 20, max_d=32, max_k=256, labeling=labeling)
    for i, (n, d, k, seed) in enumerate(clkmeans.job_mix(20, seed=3, log2n=(10, 16))):
        if d <= 32:
            p2, c2 = make_job(n, d, k, seed)
            mixed.add(p2, c2, max_iterations=3 + i % 5, stream=stream.handle)
    s = mixed.add(pts, init, max_iterations=6, stream=stream.handle)
    for _ in range(8):
        mixed.step(stream.handle)
    r_mix = mixed.take(s, stream.handle)
    mixed.close()
    for a, m in zip(r_alone[:3], r_mix[:3]):
        np.testing.assert_array_equal(a.view(np.uint32), m.view(np.uint32))
    assert r_alone[3] == r_mix[3] == 6 and r_alone[4] == r_mix[4]


def test_batch_tolerance_stops_jobs(gpu_ctx, stream):
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=64 << 20, max_d=16, max_k=64)
    slots = []
    for i in range(4):
        pts, init = make_job(20000 + 1000 * i, 8, 12, seed=30 + i)
        slots.append((b.add(pts, init, max_iterations=50, tolerance=1e-4, stream=stream.handle), pts, init))
    for _ in range(50):
        b.step(stream.handle)
        if b.poll(stream.handle)[1] == 0:
            break
    for s, pts, init in slots:
        c, l, m, it, shift = b.take(s, stream.handle)
        assert it < 50 and shift < 1e-4
        check_job(pts, init, it, (c, l, m, it, shift))
    b.close()


def test_batch_rejects_jobs_beyond_32bit_offsets(gpu_ctx, stream):
    """A job of more than 2^30 coordinates is refused with CLK_ERR_SIZE before
    anything is read or copied (np.zeros leaves its pages untouched)."""
    b = clkmeans.Batch(gpu_ctx, max_jobs=2, arena_bytes=64 << 20, max_d=32, max_k=8)
    pts = np.zeros((32, (1 << 25) + 1), np.float32)
    with pytest.raises(clkmeans.ClkError) as e:
        b.add(pts, np.zeros((32, 8), np.float32), 2, stream=stream.handle)
    assert e.value.code == clkmeans.CLK_ERR_SIZE
    del pts
    b.close()


def test_batch_admission_by_footprint(gpu_ctx, stream):
    pts, init = make_job(100000, 16, 20, seed=1)
    cfg_bytes = clkmeans.Batch(gpu_ctx, 1, 1 << 20, 16, 20).job_bytes(100000, 16, 20)
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=int(cfg_bytes * 2.5), max_d=16, max_k=20)
    s1 = b.add(pts, init, 2, stream=stream.handle)
    s2 = b.add(pts, init, 2, stream=stream.handle)
    with pytest.raises(clkmeans.ClkError) as e:
        b.add(pts, init, 2, stream=stream.handle)
    assert e.value.code == clkmeans.CLK_ERR_NOMEM
    for _ in range(2):
        b.step(stream.handle)
    r1 = b.take(s1, stream.handle)
    s3 = b.add(pts, init, 2, stream=stream.handle)  # fits again
    for _ in range(2):
        b.step(stream.handle)
    r2, r3 = b.take(s2, stream.handle), b.take(s3, stream.handle)
    for r in (r1, r2, r3):
        check_job(pts, init, 2, r)
    b.close()


def test_batch_matches_lloyd_step(gpu_ctx, stream):
    # the same job through the single-job entry point on its batch buffers
    n, d, k = 30000, 8, 25
    pts, init = make_job(n, d, k, seed=9)
    b = clkmeans.Batch(gpu_ctx, max_jobs=2, arena_bytes=64 << 20, max_d=8, max_k=32)
    s = b.add(pts, init, 5, stream=stream.handle)
    for _ in range(5):
        b.step(stream.handle)
    c_b, l_b, m_b, _, _ = b.take(s, stream.handle)
    b.close()
    dp = hip.DeviceArray.from_host(pts)
    cen = [hip.DeviceArray.from_host(init), hip.DeviceArray((d, k), np.float32)]
    sums = hip.DeviceArray((d, k), np.float64)
    masses = hip.DeviceArray((k,), np.uint32)
    labels = hip.DeviceArray((n,), np.uint32)
    cfg = gpu_ctx.config(clkmeans.STAGE_FUSED, "cluster_merge")
    for _ in range(5):
        gpu_ctx.lloyd_step(cfg, d, n, k, dp, cen[0], cen[1], sums, labels, masses, stream=stream.handle)
        cen.reverse()
    stream.synchronize()
    np.testing.assert_array_equal(labels.to_host(), l_b)
    np.testing.assert_array_equal(masses.to_host(), m_b)
    np.testing.assert_allclose(cen[0].to_host(), c_b, rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("labeling", ["mfma_fp16", "mfma_bf16"])
def test_batch_mfma_near_ties_and_special_values(gpu_ctx, stream, labeling):
    # uniform data in 32 dimensions (many points need the exact refinement),
    # lattice ties, a NaN point and huge values (exact scan)
    rng = np.random.default_rng(4)
    jobs = []
    u = rng.random((32, 6000), dtype=np.float32)
    jobs.append((u, np.ascontiguousarray(rng.random((32, 200), dtype=np.float32))))
    g = np.arange(0, 7, 2, dtype=np.float32)
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), 0).reshape(3, -1)
    p = np.arange(-1, 8, dtype=np.float32)
    lat = np.ascontiguousarray(np.stack(np.meshgrid(p, p, p, indexing="ij"), 0).reshape(3, -1))
    jobs.append((lat, np.ascontiguousarray(cen)))
    pts, init = make_job(5000, 8, 16, seed=12)
    pts = pts.copy()
    pts[:, 77] = np.nan
    pts[:, 300:364] *= np.float32(1e19)
    jobs.append((pts, init))
    b = clkmeans.Batch(gpu_ctx, max_jobs=8, arena_bytes=64 << 20, max_d=32, max_k=256, labeling=labeling)
    slots = [b.add(p_, c_, 1, stream=stream.handle) for p_, c_ in jobs]
    b.step(stream.handle)
    for s_, (p_, c_) in zip(slots, jobs):
        c, l, m, it, _ = b.take(s_, stream.handle)
        ref = O.label(np.ascontiguousarray(p_), np.ascontiguousarray(c_))
        np.testing.assert_array_equal(l, ref)
        np.testing.assert_array_equal(m, O.mass(ref, c_.shape[1]))
    b.close()


def test_batch_autotune_picks_a_tiling_and_reproduces(gpu_ctx, stream):
    jobs = [make_job(n, d, k, seed=200 + i) for i, (n, d, k) in enumerate(MIX[:8])]
    b = clkmeans.Batch(gpu_ctx, max_jobs=16, arena_bytes=256 << 20, max_d=64, max_k=256, labeling="mfma_fp16")
    slots = [b.add(p, c, 5, stream=stream.handle) for p, c in jobs]
    chosen = b.autotune(stream.handle)
    assert chosen in (4, 8, 16, 32, 64, 128, 256)
    for _ in range(5):
        b.step(stream.handle)
    tuned = [b.take(s, stream.handle) for s in slots]
    with pytest.raises(clkmeans.InvalidArgument):  # no active job to time
        b.autotune(stream.handle)
    b.close()
    # tuning between iterations (the dry passes then include the keep test)
    # changes no bit: the sums are exact
    mid = clkmeans.Batch(gpu_ctx, max_jobs=16, arena_bytes=256 << 20, max_d=64, max_k=256, labeling="mfma_fp16")
    slots = [mid.add(p, c, 5, stream=stream.handle) for p, c in jobs]
    for _ in range(2):
        mid.step(stream.handle)
    assert mid.autotune(stream.handle) in (4, 8, 16, 32, 64, 128, 256)
    for _ in range(3):
        mid.step(stream.handle)
    for s, r in zip(slots, tuned):
        r2 = mid.take(s, stream.handle)
        for a, b_ in zip(r[:3], r2[:3]):
            np.testing.assert_array_equal(a.view(np.uint32), b_.view(np.uint32))
        assert r2[3] == 5
    mid.close()
    # the same tile_work in a fresh batch: the same bits; the iteration is the reference's
    again = clkmeans.Batch(gpu_ctx, max_jobs=16, arena_bytes=256 << 20, max_d=64, max_k=256, tile_work=chosen,
                           labeling="mfma_fp16")
    slots = [again.add(p, c, 5, stream=stream.handle) for p, c in jobs]
    for _ in range(5):
        again.step(stream.handle)
    for s, (p, c), r in zip(slots, jobs, tuned):
        r2 = again.take(s, stream.handle)
        for a, b_ in zip(r[:3], r2[:3]):
            np.testing.assert_array_equal(a.view(np.uint32), b_.view(np.uint32))
        check_job(p, c, 5, r)
    again.close()
