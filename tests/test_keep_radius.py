"""The keep-radius shortcut of the bounded fused pass, checked on the CPU
against the oracle's full FP32 scan: a point whose FP32 distance to centroid
a is at most a's keep threshold must get label a from the full scan
(kmeans_naive.cpp:47-69), at every scale, dimension and for adversarial
points placed on the threshold towards the nearest other centroid.

keep_threshold below restates the device function (clk_common.hpp): s2 =
squared FP64 distance to the nearest other centroid, thr = s2 (1-m)/(2+m)^2
with m = 2^-10, rounded down to float, capped at 1e30, -1 below 1e-20.  A
planted mistake in it (the margin dropped, or the factor 1/4 dropped) must
make the check fail, so the check has teeth.
"""
import numpy as np
import pytest

from oracle import oracle as O

M = 1.0 / 1024.0


def keep_threshold(cen, a, margin=M, quarter=True):
    c = cen.astype(np.float64)
    s2 = np.inf
    for j in range(c.shape[1]):
        if j != a:
            v = float(np.sum((c[:, a] - c[:, j]) ** 2))
            if v < s2:
                s2 = v
    if not s2 > 0:
        return np.float32(-1)
    t = s2 * (1 - margin) / ((2 + margin) ** 2 if quarter else 1.0)
    t = min(t, 1e30)
    if t < 1e-20:
        return np.float32(-1)
    f = np.float32(t)
    if float(f) > t:
        f = np.nextafter(f, np.float32(0))
    return f


def boundary_points(rng, cen, a, thr):
    """Points with FP32 distance to c_a just under and over thr: towards the
    nearest other centroid, towards every other one, and in random directions."""
    d, k = cen.shape
    c = cen.astype(np.float64)
    r = np.sqrt(max(float(thr), 0.0))
    dirs = []
    for j in range(k):
        if j != a and np.isfinite(c[:, j]).all():
            v = c[:, j] - c[:, a]
            if np.linalg.norm(v) > 0:
                dirs.append(v / np.linalg.norm(v))
    for _ in range(8):
        v = rng.normal(size=d)
        dirs.append(v / np.linalg.norm(v))
    cols = []
    for u in dirs:
        for f in (1 - 1e-4, 1 - 1e-6, 1 - 1e-7, 1 - 3e-8, 1.0, 1 + 3e-8, 1 + 1e-7, 1 + 1e-6):
            cols.append(c[:, a] + u * r * f)
    return np.ascontiguousarray(np.array(cols).T.astype(np.float32))


def violations(rng, trials, **kw):
    """(#points with D(x, c_a) <= thr, #of them the full scan labels otherwise)."""
    kept = bad = 0
    for scale, d, k, close in trials:
        cen = (rng.uniform(-1, 1, size=(d, k)) * scale).astype(np.float32)
        if close:  # a pair of centroids a few ulps apart, another pair close
            cen[:, 1] = np.nextafter(cen[:, 0], np.float32(np.inf))
            cen[:, 2] = cen[:, 3] + np.float32(scale * 1e-3)
        for a in range(k):
            thr = keep_threshold(cen, a, **kw)
            if thr < 0:
                continue
            pts = boundary_points(rng, cen, a, thr)
            dist = O.distance(pts, cen, a)
            inside = dist <= thr
            kept += int(inside.sum())
            lab = O.label(pts, cen)
            bad += int((lab[inside] != a).sum())
    return kept, bad


def trials(rng, n):
    out = []
    for _ in range(n):
        close = bool(rng.integers(0, 2))
        out.append((float(10.0 ** rng.uniform(-8, 8)), int(rng.choice([1, 2, 3, 4, 8, 16, 64])),
                    int(rng.integers(4 if close else 2, 12)), close))
    return out


def test_keep_threshold_implies_full_scan_label():
    rng = np.random.default_rng(11)
    kept, bad = violations(rng, trials(rng, 120))
    assert kept > 5000  # many points sit on the threshold (not vacuous)
    assert bad == 0


@pytest.mark.parametrize("kw", [{"margin": 0.0}, {"quarter": False}])
def test_planted_threshold_mistakes_are_caught(kw):
    rng = np.random.default_rng(11)
    _, bad = violations(rng, trials(rng, 120), **kw)
    assert bad > 0


def test_keep_threshold_special_cases():
    cen = np.array([[0.0, 0.0, 3.0, np.nan]], np.float32)
    assert keep_threshold(cen, 0) == -1 and keep_threshold(cen, 1) == -1  # duplicates
    t = keep_threshold(cen, 2)  # nearest finite other at distance 3 (NaN skipped)
    assert 0 < t <= 9 / 4 and t > 9 / 4 * (1 - 4 * M)
    t1 = keep_threshold(np.array([[1.0]], np.float32), 0)  # k == 1: the cap, rounded down
    assert float(t1) <= 1e30 and t1 == np.nextafter(np.float32(1e30), np.float32(0))
