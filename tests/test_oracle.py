"""Pins the CPU oracle (oracle/kmeans_oracle.c) before it is trusted.

The reference tree holds no golden vectors for the Lloyd path, so each oracle
function is checked against something independent of its own code:
  - exact rational arithmetic (fractions) for the fma-sequence distance and
    the strict-'<' argmin, on tiny inputs with adversarial near-ties;
  - the R stats kmeans_Lloyd / kmeans_Lloyd_float from the reference's
    baselines/r_stats_kmeans.c, compiled into oracle/_ref (when present);
  - scikit-learn's Lloyd KMeans with the same initial centroids;
  - closed forms (k == 1 mean, bincount masses, fsum sums, first_x);
  - the reference's ReduceVectorParcol test shapes (test/reduce_vector.cpp).
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import clkmeans
from oracle import oracle as O

# ---------------------------------------------------------------------------
# exact-arithmetic restatement (independent of the C code)


def round_f32(q):
    """Round a rational to the nearest float32 (ties to even), as a Fraction."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    # fix e so that 2^e <= a < 2^(e+1)
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)                        # subnormal range keeps the 2^-149 quantum
    quantum = Fraction(2) ** (e - 23)
    m = a / quantum
    lo = math.floor(m)
    rem = m - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    r = lo * quantum
    if r >= Fraction(2) ** 128:
        return Fraction(sign) * Fraction(10) ** 400  # overflow marker -> inf
    return sign * r


def exact_label(points, cen):
    d, n = points.shape
    k = cen.shape[1]
    FLT_MAX = Fraction(np.finfo(np.float32).max.item())
    out = []
    for p in range(n):
        best, lab = FLT_MAX, 0
        for c in range(k):
            dist = Fraction(0)
            for f in range(d):
                t = round_f32(Fraction(float(points[f, p])) - Fraction(float(cen[f, c])))
                dist = round_f32(t * t + dist)
            if dist < best:
                best, lab = dist, c
        out.append(lab)
    return np.array(out, np.uint32)


def near_tie_case(rng, d, n, k):
    """Centroids placed so that many points are almost equidistant from two of
    them: a wrong operation order or a dropped fma flips some labels."""
    cen = rng.standard_normal((d, k)).astype(np.float32)
    pts = np.empty((d, n), np.float32)
    for p in range(n):
        a, b = rng.choice(k, 2, replace=False)
        mid = (cen[:, a].astype(np.float64) + cen[:, b]) / 2
        pts[:, p] = (mid + rng.standard_normal(d) * 1e-6).astype(np.float32)
    return pts, cen


def test_label_matches_exact_rational_on_near_ties():
    rng = np.random.default_rng(0)
    total_flips = 0
    for trial in range(25):
        d = int(rng.integers(1, 6))
        k = int(rng.integers(2, 6))
        pts, cen = near_tie_case(rng, d, 12, k)
        exact = exact_label(pts, cen)
        np.testing.assert_array_equal(O.label(pts, cen), exact)
        # the same data labelled in float64 (no fp32 rounding) disagrees
        # somewhere: the cases are sensitive to the arithmetic sequence
        diff = ((pts.astype(np.float64)[:, :, None] - cen.astype(np.float64)[:, None, :]) ** 2).sum(0)
        total_flips += int((diff.argmin(1) != exact).sum())
    assert total_flips > 0, "near-tie cases do not exercise the rounding sequence"


def test_distance_matches_exact_rational():
    # ora_dist_f32 is the value label() compares: an exact-rational fma chain
    rng = np.random.default_rng(4)
    for _ in range(10):
        d, n, k = int(rng.integers(1, 6)), 8, int(rng.integers(1, 5))
        pts, cen = near_tie_case(rng, d, n, max(k, 2))
        for c in range(cen.shape[1]):
            got = O.distance(pts, cen, c)
            for p in range(n):
                dist = Fraction(0)
                for f in range(d):
                    t = round_f32(Fraction(float(pts[f, p])) - Fraction(float(cen[f, c])))
                    dist = round_f32(t * t + dist)
                assert Fraction(float(got[p])) == dist


def test_label_random_small_exact():
    rng = np.random.default_rng(1)
    for _ in range(20):
        d, n, k = int(rng.integers(1, 5)), 20, int(rng.integers(1, 7))
        pts = (rng.standard_normal((d, n)) * 10).astype(np.float32)
        cen = (rng.standard_normal((d, k)) * 10).astype(np.float32)
        np.testing.assert_array_equal(O.label(pts, cen), exact_label(pts, cen))


def test_label_ties_nan_inf():
    pts = np.array([[0.0, 1.0, 5.0]], np.float32)
    cen = np.array([[2.0, 0.0, 2.0, np.nan]], np.float32)
    # p=1: |1-2| == |1-0| -> centroids 0 and 1 tie -> lowest index 0
    np.testing.assert_array_equal(O.label(pts, cen), [1, 0, 0])
    # all NaN -> label 0 (nothing < FLT_MAX)
    assert np.all(O.label(pts, np.full((1, 3), np.nan, np.float32)) == 0)
    # overflowing distance (inf) never beats FLT_MAX -> stays 0 unless a finite one exists
    big = np.array([[3e38]], np.float32)
    cen2 = np.array([[-3e38, 3e38]], np.float32)
    np.testing.assert_array_equal(O.label(big, cen2), [1])
    cen3 = np.array([[-3e38, -2e38]], np.float32)
    np.testing.assert_array_equal(O.label(big, cen3), [0])


def test_mass_sums_divide_closed_forms():
    rng = np.random.default_rng(2)
    pts = (rng.standard_normal((3, 5000)) * 30).astype(np.float32)
    labels = rng.integers(0, 7, 5000).astype(np.uint32)
    m = O.mass(labels, 7)
    np.testing.assert_array_equal(m, np.bincount(labels, minlength=7))
    assert m.sum() == 5000
    s64 = O.sums_f64(pts, labels, 7)
    for f in range(3):
        for c in range(7):
            exact = math.fsum(float(x) for x in pts[f, labels == c])
            assert abs(s64[f, c] - exact) <= 1e-12 * np.abs(pts[f, labels == c]).sum()
    kah = O.sums_kahan(pts, labels, 7)
    assert np.all(np.abs(kah - s64) <= 2 ** -23 * np.abs(s64) + 1e-3)
    m[4] = 0          # an empty cluster has a zero sum: 0/0 = NaN
    s64[:, 4] = 0.0
    kah[:, 4] = 0.0
    cen = O.divide_from_f64(s64, m)
    assert np.all(np.isnan(cen[:, 4]))
    live = m != 0
    np.testing.assert_array_equal(cen[:, live], (s64[:, live] / m[live]).astype(np.float32))
    cen32 = O.divide_f32(kah, m)
    np.testing.assert_array_equal(cen32[:, live], kah[:, live] / m[live].astype(np.float32))


def test_double_triple_closed_forms():
    # the double / uint64 / uint64 functions: bincount, math.fsum per cell,
    # and a Lloyd step of the float64 loop equal to those pieces
    rng = np.random.default_rng(5)
    pts = rng.standard_normal((4, 3000)) * 30
    labels = rng.integers(0, 9, 3000).astype(np.uint64)
    np.testing.assert_array_equal(O.mass_u64(labels, 9), np.bincount(labels.astype(np.int64), minlength=9))
    s = O.sums_f64_dp(pts, labels, 9)
    for f in range(4):
        for c in range(9):
            cell = pts[f, labels == c]
            assert abs(s[f, c] - math.fsum(cell)) <= 1e-12 * np.abs(cell).sum()
    init = np.ascontiguousarray(pts[:, :9])
    c1, l1, m1 = O.lloyd_f64(pts, init, 1)
    np.testing.assert_array_equal(l1, O.label_f64(pts, init))
    np.testing.assert_array_equal(m1, O.mass_u64(l1, 9))
    np.testing.assert_array_equal(c1, O.sums_f64_dp(pts, l1, 9) / m1.astype(np.float64))


def test_first_x_and_k1():
    pts, _ = clkmeans.generate_blobs(1000, 3, 4, seed=3)
    fx = O.first_x(pts, 6)
    np.testing.assert_array_equal(fx, pts[:, [0, 1, 2, 3, 4, 5]])
    c, l, m = O.lloyd(pts, pts[:, :1], 3)
    assert np.all(l == 0) and m[0] == 1000
    np.testing.assert_allclose(c[:, 0], pts.astype(np.float64).mean(1), rtol=1e-6)


@pytest.mark.skipif(not os.path.exists(O.R_REF_PATH), reason="oracle/_ref not built (reference tree absent)")
def test_against_reference_tree_r_lloyd():
    """R's float Lloyd from the reference tree (fma-contracted as the reference
    builds it): one iteration's labels equal ours exactly; double Lloyd agrees
    with our loop on separated blobs."""
    rng = np.random.default_rng(4)
    for _ in range(30):
        d, n, k = int(rng.integers(1, 6)), 200, int(rng.integers(2, 9))
        pts = (rng.standard_normal((d, n)) * 5).astype(np.float32)
        cen = (rng.standard_normal((d, k)) * 5).astype(np.float32)
        _, rl, _, _ = O.r_lloyd(pts, cen, 1)
        np.testing.assert_array_equal(O.label(pts, cen), rl)
    pts, _ = clkmeans.generate_blobs(20000, 4, 10, seed=0)
    init = O.first_x(pts, 10)
    rc, rl, rn, it = O.r_lloyd(pts.astype(np.float64), init.astype(np.float64), 30, double=True)
    oc, ol, om = O.lloyd(pts, init, 30)
    np.testing.assert_array_equal(ol, rl)
    np.testing.assert_array_equal(om, rn)
    np.testing.assert_allclose(oc, rc, rtol=1e-6)
    rc32, rl32, rn32, _ = O.r_lloyd(pts, init, 30)
    np.testing.assert_array_equal(ol, rl32)


def test_against_sklearn_lloyd():
    from sklearn.cluster import KMeans
    pts, _ = clkmeans.generate_blobs(20000, 4, 10, seed=1)
    init = O.first_x(pts, 10)
    km = KMeans(n_clusters=10, init=init.T.astype(np.float64), n_init=1, max_iter=30, tol=0.0,
                algorithm="lloyd").fit(pts.T.astype(np.float64))
    oc, ol, om = O.lloyd(pts, init, 30)
    np.testing.assert_array_equal(ol, km.labels_)
    np.testing.assert_allclose(oc.T, km.cluster_centers_, rtol=1e-5, atol=1e-5)


def test_lloyd_mt_equals_single_threaded():
    pts, _ = clkmeans.generate_blobs(50000, 4, 10, seed=2, sigma=4.0, box=10.0)
    init = pts[:, ::7][:, :10].copy()
    a = O.lloyd(pts, init, 8)
    b = O.lloyd_mt(pts, init, 8, 4)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_allclose(a[0], b[0], rtol=1e-6)


def test_naive_kahan_mode_close_to_fp64():
    pts, _ = clkmeans.generate_blobs(30000, 4, 10, seed=6)
    init = O.first_x(pts, 10)
    a = O.lloyd(pts, init, 10)
    b = O.lloyd(pts, init, 10, mode=O.MODE_NAIVE_KAHAN)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[0], b[0], rtol=1e-6)


# The reference's ReduceVectorParcol cases (test/reduce_vector.cpp:74-221).
@pytest.mark.parametrize("rows,cols,kind", [(4, 256, "ascending"), (4, 256, "large"), (4, 256, "random"),
                                            (4, 65536, "random"), (256, 46080, "random")])
def test_reduce_u32_matches_reference_verifier(rows, cols, kind):
    if kind == "ascending":
        data = np.arange(rows * cols, dtype=np.uint64).astype(np.uint32)
    elif kind == "large":
        data = np.full(rows * cols, 1 << 31, np.uint32)
    else:
        data = np.random.default_rng(cols).integers(0, 2 ** 32, rows * cols, dtype=np.uint64).astype(np.uint32)
    ref = (data.reshape(cols, rows).astype(np.uint64).sum(0) % (1 << 32)).astype(np.uint32)
    np.testing.assert_array_equal(O.reduce_u32(data, cols, rows), ref)


def test_partition_matrix_layout():
    """buffer_helper.cpp:58-84: chunk b holds, for each of num_dims rows, the
    row's bytes [b*bs/num_dims, ...) packed back to back."""
    d, n, item = 3, 1000, 4
    src = np.arange(d * n * item, dtype=np.uint32).astype(np.uint8).tobytes()
    buffer_size = 3 * 4 * 64  # 64 points per chunk
    nb, dst = O.partition_matrix(src, d, buffer_size)
    s = np.frombuffer(src, np.uint8).reshape(d, n * item)
    per = buffer_size // d
    assert nb == -(-len(src) // buffer_size)
    for b in range(nb):
        real = min(per, n * item - b * per)
        for v in range(d):
            got = dst[b * buffer_size + v * real: b * buffer_size + (v + 1) * real]
            np.testing.assert_array_equal(got, s[v, b * per: b * per + real])
