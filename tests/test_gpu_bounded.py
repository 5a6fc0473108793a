"""The bounded fused pass ("cluster_merge_bounded"): a point keeps its
previous label after one distance when that distance is inside the label's
keep radius, and scans every centroid otherwise.  Whatever the labels buffer
holds on entry, labels, masses, FP64 sums and divided centroids must be
bit-identical to the exhaustive "cluster_merge" pass, and labels equal to the
oracle's full scan (kmeans_naive.cpp:47-69).
"""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

FUSED = clkmeans.STAGE_FUSED
M = 1.0 / 1024.0


def keep_radius2(cen, a):
    """Squared keep radius of centroid a (the kernel's threshold, before the
    float rounding): s2 (1-m) / (2+m)^2, s2 = squared distance to the nearest
    other centroid.  Used here only to place points at the boundary."""
    c = cen.astype(np.float64)
    s2 = np.array([np.sum((c[:, a] - c[:, j]) ** 2) for j in range(c.shape[1]) if j != a])
    s2 = s2[~np.isnan(s2)]
    return s2.min() * (1 - M) / (2 + M) ** 2


class Pass:
    """One fused Lloyd step (clk_lloyd_step) with its own device buffers."""

    def __init__(self, ctx, strategy, pts, init, labels0, **kw):
        self.ctx = ctx
        self.d, self.n = pts.shape
        self.k = init.shape[1]
        self.cfg = ctx.config(FUSED, strategy, **kw)
        self.dp = dev(pts)
        self.cen = [dev(init), zeros(init.shape, np.float32)]
        self.sums = zeros(init.shape, np.float64)
        self.masses = zeros((self.k,), np.uint32)
        self.labels = dev(labels0)

    def step(self):
        old, new = self.cen
        self.ctx.lloyd_step(self.cfg, self.d, self.n, self.k, self.dp, old, new, self.sums, self.labels,
                            self.masses)
        self.cen = [new, old]
        hip.synchronize()
        return (self.labels.to_host()[:self.n], self.masses.to_host(), self.sums.to_host(),
                self.cen[0].to_host())


def assert_same(a, b):
    la, ma, sa, ca = a
    lb, mb, sb, cb = b
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(sa.view(np.uint64), sb.view(np.uint64))
    np.testing.assert_array_equal(ca.view(np.uint32), cb.view(np.uint32))


LLOYD_CASES = [
    # n, d, k, sigma, box, config overrides
    (1 << 16, 4, 10, 1.0, 50.0, {}),
    ((1 << 14) + 2, 4, 10, 1.0, 50.0, {}),                 # two points per work-item
    (777, 5, 4, 1.0, 10.0, {}),                            # one point per work-item, padded d
    (40000, 3, 7, 2.0, 10.0, {}),                          # overlapping blobs: many full scans
    (30000, 1, 5, 1.0, 20.0, {}),
    (20000, 2, 1, 1.0, 20.0, {}),                          # k == 1
    (30000, 8, 12, 1.0, 20.0, {}),
    (10000, 16, 6, 1.0, 20.0, {"local_size": 128}),
    (4096, 32, 3, 1.0, 20.0, {}),
    (2048, 64, 2, 1.0, 20.0, {}),
    (20000, 4, 10, 1.0, 50.0, {"global_size": 4096, "local_size": 32}),
]


@pytest.mark.parametrize("n,d,k,sigma,box,kw", LLOYD_CASES)
def test_bounded_lloyd_equals_exhaustive(gpu_ctx, n, d, k, sigma, box, kw):
    # (every case fits the private tile, k (d+1) columns of local_size
    # floats in LDS: its FP32 sums have a fixed order, so bits must match)
    assert k * (d + 1) * kw.get("local_size", 256) * 4 <= 160 * 1024
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=sigma, box=box, seed=n + d + k)
    init = O.first_x(pts, k)
    rng = np.random.default_rng(7)
    garbage = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)  # mostly >= k
    full = Pass(gpu_ctx, "cluster_merge", pts, init, np.zeros(n, np.uint32), **kw)
    bnd = Pass(gpu_ctx, "cluster_merge_bounded", pts, init, garbage, **kw)
    cen = init
    for it in range(8):
        a, b = full.step(), bnd.step()
        assert_same(a, b)
        if it in (0, 7):
            np.testing.assert_array_equal(b[0], O.label(pts, cen))
        cen = b[3]


def test_bounded_keep_boundary(gpu_ctx):
    # Points at the keep radius of their previous label, just inside and
    # just outside, towards the nearest other centroid (the direction where
    # the keep test is tightest), at the midpoints (ties: the lower index
    # wins), at the centroids themselves, and random; centroid 3 duplicates
    # centroid 1 (no keep radius), centroid 4 is NaN (an empty cluster) and
    # centroid 5 has an infinite coordinate.
    rng = np.random.default_rng(3)
    d, k = 4, 8
    cen = rng.uniform(-5, 5, size=(d, k)).astype(np.float32)
    cen[:, 2] = cen[:, 0] + np.float32(0.25)
    cen[:, 3] = cen[:, 1]
    cen[:, 4] = np.nan
    cen[0, 5] = np.inf
    c64 = cen.astype(np.float64)
    cols, prev = [], []
    for a in (0, 1, 2, 6, 7):
        r = np.sqrt(keep_radius2(cen, a))
        # the nearest other finite centroid at a nonzero distance (centroid
        # 1's duplicate, centroid 3, gives no direction)
        others = [j for j in range(k) if j != a and np.isfinite(c64[:, j]).all()
                  and np.sum((c64[:, a] - c64[:, j]) ** 2) > 0]
        near = min(others, key=lambda j: np.sum((c64[:, a] - c64[:, j]) ** 2))
        u = (c64[:, near] - c64[:, a]) / np.linalg.norm(c64[:, near] - c64[:, a])
        # a keep radius of 0 (a duplicated centroid: never keep) has no
        # boundary; the points along u then lie at the centroid
        rb = r if r > 0 else np.sqrt(np.sum((c64[:, near] - c64[:, a]) ** 2)) / 2
        for f in (1 - 1e-6, 1 - 1e-7, 1.0, 1 + 1e-7, 1 + 1e-6, 1.01, 2.2):
            cols.append(c64[:, a] + u * rb * f)
            prev.append(a)
        cols.append(0.5 * (c64[:, a] + c64[:, near]))  # equidistant (in real arithmetic)
        prev.append(max(a, near))
        cols.append(c64[:, a])
        prev.append(a)
        for _ in range(40):
            v = rng.normal(size=d)
            cols.append(c64[:, a] + v / np.linalg.norm(v) * r * rng.uniform(0.9, 1.1))
            prev.append(a)
    pts = np.ascontiguousarray(np.array(cols).T.astype(np.float32))
    assert np.isfinite(pts).all()  # NaN / inf only in the centroids
    n = pts.shape[1]
    prev = np.array(prev, np.uint32)
    ref = O.label(pts, cen)
    for start in (prev, ref, np.full(n, 4, np.uint32), np.full(n, 3, np.uint32)):
        outs = []
        for strategy in ("cluster_merge", "cluster_merge_bounded"):
            dl = dev(start)
            ds, dm = zeros((d, k), np.float64), zeros((k,), np.uint32)
            gpu_ctx.fused(gpu_ctx.config(FUSED, strategy, vector_length=1), d, n, k, dev(pts), dev(cen), ds,
                          dl, dm)
            hip.synchronize()
            outs.append((dl.to_host(), dm.to_host(), ds.to_host()))
        np.testing.assert_array_equal(outs[1][0], ref)
        np.testing.assert_array_equal(outs[0][0], outs[1][0])
        np.testing.assert_array_equal(outs[0][1], outs[1][1])
        np.testing.assert_array_equal(outs[0][2].view(np.uint64), outs[1][2].view(np.uint64))


def test_bounded_accumulates_like_cluster_merge(gpu_ctx):
    # clk_fused (accumulate semantics, as the buffered strategies call it per
    # chunk) with a labels buffer from another chunk: same results.
    pts, _ = clkmeans.generate_blobs(50000, 4, 10, sigma=1.0, box=50.0, seed=9)
    cen = O.first_x(pts, 10)
    d, n = pts.shape
    other = O.label(np.ascontiguousarray(pts[:, ::-1]), cen)
    outs = []
    for strategy in ("cluster_merge", "cluster_merge_bounded"):
        ds = dev(np.full((d, 10), 0.5))
        dm = dev(np.arange(10, dtype=np.uint32))
        dl = dev(other)
        gpu_ctx.fused(gpu_ctx.config(FUSED, strategy), d, n, 10, dev(pts), dev(cen), ds, dl, dm)
        hip.synchronize()
        outs.append((dl.to_host(), dm.to_host(), ds.to_host()))
    np.testing.assert_array_equal(outs[1][0], O.label(pts, cen))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


def test_bounded_beyond_private_tile(gpu_ctx):
    # k (d+1) columns of 256 floats beyond LDS: the pass runs the shared-tile
    # kernel (LDS float atomics, no fixed summation order); labels and masses
    # stay exact against the oracle at every iteration.
    n, d, k = 30000, 8, 20
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=20.0, seed=5)
    cen = O.first_x(pts, k)
    bnd = Pass(gpu_ctx, "cluster_merge_bounded", pts, cen, np.zeros(n, np.uint32))
    for _ in range(4):
        lab, mass, _, nxt = bnd.step()
        ref = O.label(pts, cen)
        np.testing.assert_array_equal(lab, ref)
        np.testing.assert_array_equal(mass, np.bincount(ref, minlength=k).astype(np.uint32))
        cen = nxt


LABEL_CASES = [
    # n, d, k, config overrides
    (1 << 16, 4, 10, {}),
    (777, 5, 4, {}),
    (30000, 8, 12, {"local_size": 128}),
    (4096, 64, 3, {}),
    (3000, 16, 1200, {}),        # centroids beyond LDS: the chunked scan (no keep test)
]


@pytest.mark.parametrize("n,d,k,kw", LABEL_CASES)
def test_bounded_labeling_equals_scan(gpu_ctx, n, d, k, kw):
    # "unroll_vector_bounded" keeps the labels buffer between calls as the
    # three-stage strategies do; garbage first, then the previous labels
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=20.0, seed=n + k)
    cen = O.first_x(pts, k)
    rng = np.random.default_rng(2)
    dl = dev(rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32))
    cfg = gpu_ctx.config(clkmeans.STAGE_LABELING, "unroll_vector_bounded", **kw)
    dp = dev(pts)
    for _ in range(4):
        gpu_ctx.labeling(cfg, d, n, k, dp, dev(cen), dl)
        hip.synchronize()
        ref = O.label(pts, cen)
        np.testing.assert_array_equal(dl.to_host(), ref)
        cen = O.lloyd(pts, cen, 1)[0]
