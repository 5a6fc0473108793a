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
    for b in (O.lloyd_mt(pts, init, 8, 4), O.lloyd3_mt(pts, init, 8, 3)):
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


# ---------------------------------------------------------------------------
# The double / uint64 triple: ora_label_f64 pinned independently of its code


def exact_label_f64(points, cen, fma=True, strict=True):
    """Exact-rational restatement of the FP64 labeling: t = RN(x - c),
    dist = RN(t*t + dist) (one rounding: the fma) from 0, strict '<' against
    DBL_MAX, label 0 when nothing is smaller.  fma=False / strict=False are
    the planted mistakes the checks below must catch."""
    d, n = points.shape
    k = cen.shape[1]
    dbl_max = Fraction(np.finfo(np.float64).max.item())
    rnd = lambda q: Fraction(float(q))  # Fraction -> float rounds to nearest even
    out = []
    for p in range(n):
        best, lab = dbl_max, 0
        for c in range(k):
            dist = Fraction(0)
            for f in range(d):
                t = rnd(Fraction(float(points[f, p])) - Fraction(float(cen[f, c])))
                dist = rnd(t * t + dist) if fma else rnd(rnd(t * t) + dist)
            if dist < best or (not strict and dist == best):
                best, lab = dist, c
        out.append(lab)
    return np.array(out, np.uint64)


def near_tie_case_f64(rng, d, n, k):
    """Points on (or within ~1e-16 relative of) the bisector of two centroids,
    and a duplicated centroid (exact ties): the rounding sequence and the tie
    rule decide their labels."""
    cen = rng.standard_normal((d, k))
    cen[:, k - 1] = cen[:, 0]  # exact duplicate: ties between 0 and k-1
    pts = np.empty((d, n))
    for p in range(n):
        a, b = rng.choice(k, 2, replace=False)
        mid = (cen[:, a] + cen[:, b]) / 2  # exact: equal true distances
        noise = 0.0 if p % 2 == 0 else 1e-16
        pts[:, p] = mid + rng.standard_normal(d) * noise * (1 + np.abs(mid))
    return pts, cen


def _check_label_f64(fn, seeds=range(30)):
    """The checks that pin ora_label_f64: near-tie cases against the exact
    restatement, plus R's double kmeans_Lloyd (reference tree) after one
    iteration on random data with duplicated centroids."""
    for s in seeds:
        rng = np.random.default_rng(1000 + s)
        d, k = int(rng.integers(1, 5)), int(rng.integers(3, 6))
        pts, cen = near_tie_case_f64(rng, d, 10, k)
        if not np.array_equal(fn(pts, cen), exact_label_f64(pts, cen)):
            return False
    if O.r_ref() is not None:
        for s in seeds:
            rng = np.random.default_rng(2000 + s)
            d, n, k = int(rng.integers(1, 6)), 150, int(rng.integers(2, 9))
            pts = rng.standard_normal((d, n)) * 5
            cen = rng.standard_normal((d, k)) * 5
            cen[:, -1] = cen[:, 0]
            pts[:, :k] = cen  # points on duplicated centroids: exact ties
            _, rl, _, _ = O.r_lloyd(pts, cen, 1, double=True)
            if not np.array_equal(fn(pts, cen), rl.astype(np.uint64)):
                return False
    return True


def test_label_f64_pinned_by_exact_rationals_and_r():
    assert _check_label_f64(O.label_f64)


def test_label_f64_checks_catch_planted_mistakes():
    # the same checks run on the exact restatement with a planted '<=' and
    # with the fma split into a rounded product and a rounded sum: both fail
    assert not _check_label_f64(lambda p, c: exact_label_f64(p, c, strict=False), seeds=range(8))
    assert not _check_label_f64(lambda p, c: exact_label_f64(p, c, fma=False), seeds=range(30))


@pytest.mark.skipif(not os.path.exists(O.R_REF_PATH), reason="oracle/_ref not built (reference tree absent)")
def test_lloyd_f64_against_r_double_lloyd():
    """ora_lloyd_f64 against R's double kmeans_Lloyd from the reference tree:
    labels and masses equal over 30 iterations, centroids to FP64 rounding."""
    pts, _ = clkmeans.generate_blobs(20000, 5, 8, seed=11, sigma=2.0, box=10.0)
    p64 = pts.astype(np.float64)
    init = np.ascontiguousarray(p64[:, :8])
    rc, rl, rn, _ = O.r_lloyd(p64, init, 30, double=True)
    oc, ol, om = O.lloyd_f64(p64, init, 30)
    np.testing.assert_array_equal(ol, rl.astype(np.uint64))
    np.testing.assert_array_equal(om, rn.astype(np.uint64))
    np.testing.assert_allclose(oc, rc, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# Exact grid sums (the library's summation): pinned by their definition in
# exact rational arithmetic


def grid_scales_by_definition(points, n_grid):
    b = min(25, 53 - int(n_grid).bit_length())
    out = []
    for row in np.asarray(points, np.float32):
        fin = np.abs(row[np.isfinite(row)]).astype(np.float64)
        m = Fraction(float(fin.max())) if fin.size else Fraction(0)
        if m < Fraction(2) ** -126:
            e = -100
        else:
            e = -100
            while Fraction(2) ** e <= m:  # least e with m < 2^e
                e += 1
        out.append(Fraction(2) ** (b - max(e, -100)))
    return out


def round_half_even(q):
    lo = math.floor(q)
    rem = q - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return lo


def grid_sums_exact(points, labels, k, scales, trunc=False):
    d, n = points.shape
    out = np.zeros((d, k))
    for f in range(d):
        sc = Fraction(float(scales[f]))
        for c in range(k):
            vals = points[f, labels == c]
            tot = 0
            for x in vals:
                y = Fraction(float(x)) * sc
                tot += int(y) if trunc else round_half_even(y)
            out[f, c] = float(Fraction(tot) / sc)
    return out


def test_grid_scales_match_definition():
    rng = np.random.default_rng(31)
    cases = []
    for e in (-140, -60, -3, 0, 5, 40, 120):
        cases.append(((rng.standard_normal((3, 50)) * 2.0 ** e).astype(np.float32), 50))
    z = np.zeros((2, 9), np.float32)
    z[1, 3] = 1e-45  # subnormal maximum
    cases.append((z, 9))
    sp = rng.standard_normal((2, 40)).astype(np.float32)
    sp[0, 5], sp[1, 7], sp[1, 8] = np.nan, np.inf, -np.inf  # ignored by the maximum
    cases.append((sp, 40))
    pw = np.array([[1.0, -0.5, 0.25], [4.0, 3.9999998, 2.0]], np.float32)  # maxima on powers of two
    cases.append((pw, 3))
    for n_grid in (1, 2 ** 27, 2 ** 28 - 1, 2 ** 28, 2 ** 31):
        cases.append((sp, n_grid))
    for pts, n_grid in cases:
        got = O.grid_scales(pts, n_grid)
        want = grid_scales_by_definition(pts, n_grid)
        assert [Fraction(float(g)) for g in got] == want


def test_grid_sums_match_exact_integers_and_bound():
    rng = np.random.default_rng(32)
    for trial in range(12):
        d, n, k = int(rng.integers(1, 4)), int(rng.integers(50, 400)), int(rng.integers(1, 6))
        pts = (rng.standard_normal((d, n)) * 10.0 ** rng.integers(-3, 4)).astype(np.float32)
        sc = O.grid_scales(pts)
        # values exactly half a unit off the grid: round-half-even decides them
        for f in range(d):
            unit = 1.0 / float(sc[f])
            pts[f, :8] = (np.arange(8) - 3.5) * unit * 3.0
        labels = rng.integers(0, k + 1, n).astype(np.uint32)  # label k: ignored
        got = O.sums_grid(pts, labels, k, sc)
        want = grid_sums_exact(pts, labels, k, sc)
        np.testing.assert_array_equal(got, want)
        for f in range(d):
            unit = 1.0 / float(sc[f])
            for c in range(k):
                cell = pts[f, labels == c]
                assert abs(got[f, c] - math.fsum(float(x) for x in cell)) <= cell.size * unit / 2


def test_grid_sums_special_values_and_planted_mistakes():
    pts = np.array([[1.0, np.nan, 2.0, 3.0, np.inf, -np.inf, 5.0, np.inf]], np.float32)
    labels = np.array([0, 0, 1, 2, 2, 3, 3, 4], np.uint32)
    sc = O.grid_scales(pts)
    s = O.sums_grid(pts, labels, 5, sc)[0]
    assert np.isnan(s[0]) and s[1] == 2.0 and s[2] == np.inf and s[3] == -np.inf and s[4] == np.inf
    both = np.array([[np.inf, -np.inf]], np.float32)
    assert np.isnan(O.sums_grid(both, np.zeros(2, np.uint32), 1, np.ones(1, np.float32))[0, 0])
    # a finite value outside the grid it is summed on: NaN
    assert np.isnan(O.sums_grid(np.array([[1e6]], np.float32), np.zeros(1, np.uint32), 1,
                                np.array([2.0 ** 30], np.float32))[0, 0])
    # planted mistakes: truncation instead of round-half-even, and a grid one
    # exponent too coarse, give other sums on the same data
    rng = np.random.default_rng(33)
    pts = rng.standard_normal((2, 300)).astype(np.float32)
    labels = rng.integers(0, 3, 300).astype(np.uint32)
    sc = O.grid_scales(pts)
    got = O.sums_grid(pts, labels, 3, sc)
    assert not np.array_equal(got, grid_sums_exact(pts, labels, 3, sc, trunc=True))
    assert not np.array_equal(got, grid_sums_exact(pts, labels, 3, sc / 2))


def test_grid_lloyd_close_to_fp64_merge_and_order_free():
    pts, _ = clkmeans.generate_blobs(30000, 4, 10, seed=6)
    init = O.first_x(pts, 10)
    a = O.lloyd(pts, init, 10)
    g = O.lloyd(pts, init, 10, mode=O.MODE_GRID)
    np.testing.assert_array_equal(a[1], g[1])
    np.testing.assert_allclose(a[0], g[0], rtol=1e-6)
    # the sums of a permutation of the points are the same bits
    perm = np.random.default_rng(7).permutation(pts.shape[1])
    sc = O.grid_scales(pts)
    np.testing.assert_array_equal(O.sums_grid(pts, g[1], 10, sc),
                                  O.sums_grid(pts[:, perm], g[1][perm], 10, sc))
