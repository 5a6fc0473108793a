"""Streaming passes at any n and any 4-byte aligned buffers.

Feature row f starts at f*n floats, so for n % 4 != 0 the rows are not
16-byte aligned; the passes still take 2 or 4 points per work-item step
(64/128-bit accesses at 4-byte alignment) and finish the n % width tail
points one at a time.  Here the points and labels also start one element
past a 256-byte boundary.  Every result must equal the oracle's bit for bit
(labels, masses, exact grid sums)."""
import numpy as np
import pytest

import clkmeans
from clkmeans import hip
from oracle import oracle as O
from tests.helpers import dev, zeros

pytestmark = pytest.mark.gpu

NS = [30001, 30002, 30003, 3]  # n % 4 = 1, 2, 3; fewer points than one step


class Shifted:
    """Device copies of pts [d][n] and a label array, each starting `off`
    elements into its allocation."""

    def __init__(self, pts, off, labels=None):
        d, n = pts.shape
        self.pbuf = dev(np.concatenate([np.full(off, 7.0, np.float32), pts.reshape(-1)]))
        lab = np.zeros(n, np.uint32) if labels is None else labels
        self.lbuf = dev(np.concatenate([np.full(off, 0xDEAD, np.uint32), lab]))
        self.off = off
        self.n = n
        self.points = self.pbuf.offset(off)
        self.labels = self.lbuf.offset(off)

    def labels_host(self):
        full = self.lbuf.to_host()
        assert (full[:self.off] == 0xDEAD).all(), "a label was written before the buffer"
        return full[self.off:]


def blobs(n, d, k, seed):
    if n < 16:
        rng = np.random.default_rng(seed)
        return (rng.standard_normal((d, n)) * 5).astype(np.float32)
    return clkmeans.generate_blobs(n, d, k, sigma=2.0, box=10.0, seed=seed)[0]


@pytest.mark.parametrize("off", [0, 1])
@pytest.mark.parametrize("n", NS)
@pytest.mark.parametrize("strategy,d,k,vec", [("cluster_merge", 4, 10, 4), ("cluster_merge_bounded", 4, 10, 4),
                                              ("feature_sum", 3, 7, 4), ("cluster_merge", 16, 20, 2),
                                              ("cluster_merge_bounded", 8, 33, 4)])
def test_fused(gpu_ctx, strategy, d, k, vec, n, off):
    pts = blobs(n, d, k, seed=n + d)
    k = min(k, n)
    init = O.first_x(pts, k)
    buf = Shifted(pts, off)
    dc = dev(init)
    sums, masses, newc = zeros((d, k), np.float64), zeros(k, np.uint32), zeros((d, k), np.float32)
    cfg = gpu_ctx.config(clkmeans.STAGE_FUSED, strategy, vector_length=vec)
    gpu_ctx.lloyd_step(cfg, d, n, k, buf.points, dc, newc, sums, buf.labels, masses)
    hip.synchronize()
    lab = buf.labels_host()
    np.testing.assert_array_equal(lab, O.label(pts, init))
    m = O.mass(lab, k)
    np.testing.assert_array_equal(masses.to_host(), m)
    ref = O.sums_grid(pts, lab, k, O.grid_scales(pts))
    np.testing.assert_array_equal(sums.to_host(), ref)
    with np.errstate(invalid="ignore", divide="ignore"):
        np.testing.assert_array_equal(newc.to_host(), O.divide_from_f64(ref, m))
    # the bounded pass again from these labels: the keep test reads them
    if strategy == "cluster_merge_bounded":
        gpu_ctx.lloyd_step(cfg, d, n, k, buf.points, newc, dc, sums, buf.labels, masses)
        hip.synchronize()
        np.testing.assert_array_equal(buf.labels_host(), O.label(pts, newc.to_host()))


@pytest.mark.parametrize("off", [0, 1])
@pytest.mark.parametrize("n", NS)
@pytest.mark.parametrize("strategy,d,k", [("unroll_vector", 4, 10), ("unroll_vector_bounded", 4, 10),
                                          ("unroll_vector", 32, 1400)])  # the last: centroid chunks
def test_labeling(gpu_ctx, strategy, d, k, n, off):
    pts = blobs(n, d, min(k, 64), seed=n + k)
    k = min(k, n)
    cen = O.first_x(pts, k)
    prev = O.label(pts, O.first_x(pts[:, ::-1].copy(), k))  # other labels for the keep test
    buf = Shifted(pts, off, prev.astype(np.uint32))
    cfg = gpu_ctx.config(clkmeans.STAGE_LABELING, strategy, vector_length=4)
    gpu_ctx.labeling(cfg, d, n, k, buf.points, dev(cen), buf.labels)
    hip.synchronize()
    np.testing.assert_array_equal(buf.labels_host(), O.label(pts, cen))


@pytest.mark.parametrize("off", [0, 1])
@pytest.mark.parametrize("n", NS)
def test_centroid_and_mass(gpu_ctx, n, off):
    d, k = 5, 9
    rng = np.random.default_rng(n)
    pts = (rng.standard_normal((d, n)) * 3).astype(np.float32)
    lab = rng.integers(0, k, n).astype(np.uint32)
    buf = Shifted(pts, off, lab)
    sums = zeros((d, k), np.float64)
    cfg = gpu_ctx.config(clkmeans.STAGE_CENTROID_UPDATE, "cluster_merge", vector_length=4)
    gpu_ctx.centroid_update(cfg, d, n, k, buf.points, sums, buf.labels)
    hip.synchronize()
    np.testing.assert_array_equal(sums.to_host(), O.sums_grid(pts, lab, k, O.grid_scales(pts)))
    for strategy in ("global_atomic", "part_global", "part_local", "part_private"):
        masses = zeros(k, np.uint32)
        mcfg = gpu_ctx.config(clkmeans.STAGE_MASS_UPDATE, strategy, vector_length=4)
        gpu_ctx.mass_update(mcfg, n, k, buf.labels, masses)
        hip.synchronize()
        np.testing.assert_array_equal(masses.to_host(), O.mass(lab, k), err_msg=strategy)
    # the sum grid's max pass reads the same misaligned rows
    units = gpu_ctx.sum_grid_units(d, n, buf.points)
    np.testing.assert_array_equal(units, 1.0 / O.grid_scales(pts).astype(np.float64))
