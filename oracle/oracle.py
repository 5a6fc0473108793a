"""numpy front end of the CPU oracle (liboracle.so, built from kmeans_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / timed CPU baseline.  The
product never imports it.  Reference citations live in kmeans_oracle.c.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
R_REF_PATH = os.path.join(HERE, "_ref", "libr_stats_kmeans.so")

MODE_FP64_MERGE = 0
MODE_NAIVE_KAHAN = 1
MODE_GRID = 2

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle/liboracle.so not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        sig = {
            "ora_label_f32": [vp, u64, u64, vp, u64, vp],
            "ora_label_f64": [vp, u64, u64, vp, u64, vp],
            "ora_dist_f32": [vp, u64, u64, vp, u64, u64, vp],
            "ora_mass": [vp, u64, u64, vp],
            "ora_mass_u64": [vp, u64, u64, vp],
            "ora_sum_f64_dp": [vp, u64, u64, vp, u64, vp],
            "ora_sum_f64": [vp, u64, u64, vp, u64, vp],
            "ora_sum_kahan_f32": [vp, u64, u64, vp, u64, vp],
            "ora_divide_from_f64": [vp, vp, u64, u64, vp],
            "ora_divide_f32": [vp, vp, u64, u64, vp],
            "ora_first_x": [vp, u64, u64, u64, vp],
            "ora_lloyd_f32": [vp, u64, u64, u64, vp, ctypes.c_uint32, i32, vp, vp],
            "ora_lloyd_f64": [vp, u64, u64, u64, vp, ctypes.c_uint32, vp, vp],
            "ora_lloyd_mt_f32": [vp, u64, u64, u64, vp, ctypes.c_uint32, vp, vp, i32],
            "ora_reduce_u32": [vp, u64, u64, vp],
            "ora_grid_scales": [vp, u64, u64, u64, vp],
            "ora_lloyd3_mt_f32": [vp, u64, u64, u64, vp, ctypes.c_uint32, vp, vp, i32],
            "ora_sum_grid": [vp, u64, u64, vp, u64, vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        L.ora_lloyd_mt_f32.restype = i32
        L.ora_lloyd3_mt_f32.restype = i32
        L.ora_grid_bits.argtypes = [u64]
        L.ora_grid_bits.restype = ctypes.c_uint32
        L.ora_partition_matrix.argtypes = [vp, vp, u64, u64, u64]
        L.ora_partition_matrix.restype = ctypes.c_int64
        _L = L
    return _L


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


def label(points, centroids):
    """points [d][n] float32, centroids [d][k] float32 -> uint32 labels [n]."""
    p, pp = _c(points, np.float32)
    c, cp = _c(centroids, np.float32)
    d, n = p.shape
    k = c.shape[1]
    out = np.empty(n, np.uint32)
    lib().ora_label_f32(pp, n, d, cp, k, out.ctypes.data)
    return out


def distance(points, centroids, c):
    """FP32 distance of every point to centroid c, as label() computes it."""
    p, pp = _c(points, np.float32)
    cc, cp = _c(centroids, np.float32)
    d, n = p.shape
    out = np.empty(n, np.float32)
    lib().ora_dist_f32(pp, n, d, cp, cc.shape[1], c, out.ctypes.data)
    return out


def label_f64(points, centroids):
    p, pp = _c(points, np.float64)
    c, cp = _c(centroids, np.float64)
    d, n = p.shape
    k = c.shape[1]
    out = np.empty(n, np.uint64)
    lib().ora_label_f64(pp, n, d, cp, k, out.ctypes.data)
    return out


def mass(labels, k):
    l, lp = _c(labels, np.uint32)
    out = np.empty(k, np.uint32)
    lib().ora_mass(lp, l.size, k, out.ctypes.data)
    return out


def mass_u64(labels, k):
    """uint64 masses of uint64 labels (the double / uint64 triple)."""
    l, lp = _c(labels, np.uint64)
    out = np.empty(k, np.uint64)
    lib().ora_mass_u64(lp, l.size, k, out.ctypes.data)
    return out


def sums_f64_dp(points, labels, k):
    """FP64 sums in point order of float64 points [d][n] with uint64 labels."""
    p, pp = _c(points, np.float64)
    l, lp = _c(labels, np.uint64)
    d, n = p.shape
    out = np.empty((d, k), np.float64)
    lib().ora_sum_f64_dp(pp, n, d, lp, k, out.ctypes.data)
    return out


def sums_f64(points, labels, k):
    p, pp = _c(points, np.float32)
    l, lp = _c(labels, np.uint32)
    d, n = p.shape
    out = np.empty((d, k), np.float64)
    lib().ora_sum_f64(pp, n, d, lp, k, out.ctypes.data)
    return out


def sums_kahan(points, labels, k):
    p, pp = _c(points, np.float32)
    l, lp = _c(labels, np.uint32)
    d, n = p.shape
    out = np.empty((d, k), np.float32)
    lib().ora_sum_kahan_f32(pp, n, d, lp, k, out.ctypes.data)
    return out


def grid_bits(n):
    return int(lib().ora_grid_bits(n))


def grid_scales(points, n_grid=None):
    """Per-feature grid scales 2^(b - E_f) of points [d][n] (b from n_grid,
    default n); the units are their reciprocals."""
    p, pp = _c(points, np.float32)
    d, n = p.shape
    out = np.empty(d, np.float32)
    lib().ora_grid_scales(pp, n, d, n if n_grid is None else n_grid, out.ctypes.data)
    return out


def sums_grid(points, labels, k, scales=None):
    """Exact grid sums [d][k] (FP64) of float32 points with uint32 labels."""
    p, pp = _c(points, np.float32)
    l, lp = _c(labels, np.uint32)
    d, n = p.shape
    sc = grid_scales(p) if scales is None else np.ascontiguousarray(scales, np.float32)
    out = np.empty((d, k), np.float64)
    lib().ora_sum_grid(pp, n, d, lp, k, sc.ctypes.data, out.ctypes.data)
    return out


def divide_from_f64(sums, masses):
    s, sp = _c(sums, np.float64)
    m, mp = _c(masses, np.uint32)
    d, k = s.shape
    out = np.empty((d, k), np.float32)
    lib().ora_divide_from_f64(sp, mp, d, k, out.ctypes.data)
    return out


def divide_f32(sums, masses):
    s, sp = _c(sums, np.float32)
    m, mp = _c(masses, np.uint32)
    d, k = s.shape
    out = np.empty((d, k), np.float32)
    lib().ora_divide_f32(sp, mp, d, k, out.ctypes.data)
    return out


def first_x(points, k):
    p, pp = _c(points, np.float32)
    d, n = p.shape
    out = np.empty((d, k), np.float32)
    lib().ora_first_x(pp, n, d, k, out.ctypes.data)
    return out


def lloyd(points, centroids, iterations, mode=MODE_FP64_MERGE):
    """Fixed-iteration Lloyd; returns (centroids [d][k], labels, masses)."""
    p, pp = _c(points, np.float32)
    c = np.array(centroids, dtype=np.float32, copy=True, order="C")
    d, n = p.shape
    k = c.shape[1]
    labels = np.zeros(n, np.uint32)
    masses = np.zeros(k, np.uint32)
    lib().ora_lloyd_f32(pp, n, d, k, c.ctypes.data, iterations, mode, labels.ctypes.data, masses.ctypes.data)
    return c, labels, masses


def lloyd_f64(points, centroids, iterations):
    p, pp = _c(points, np.float64)
    c = np.array(centroids, dtype=np.float64, copy=True, order="C")
    d, n = p.shape
    k = c.shape[1]
    labels = np.zeros(n, np.uint64)
    masses = np.zeros(k, np.uint64)
    lib().ora_lloyd_f64(pp, n, d, k, c.ctypes.data, iterations, labels.ctypes.data, masses.ctypes.data)
    return c, labels, masses


def lloyd_mt(points, centroids, iterations, threads):
    """Multi-threaded CPU Lloyd (the timed CPU baseline); FP64-merge semantics."""
    p, pp = _c(points, np.float32)
    c = np.array(centroids, dtype=np.float32, copy=True, order="C")
    d, n = p.shape
    k = c.shape[1]
    labels = np.zeros(n, np.uint32)
    masses = np.zeros(k, np.uint32)
    rc = lib().ora_lloyd_mt_f32(pp, n, d, k, c.ctypes.data, iterations, labels.ctypes.data,
                                masses.ctypes.data, int(threads))
    assert rc == 0
    return c, labels, masses


def lloyd3_mt(points, centroids, iterations, threads):
    """Multi-threaded three-stage CPU Lloyd (label, mass, centroid passes)."""
    p, pp = _c(points, np.float32)
    c = np.array(centroids, dtype=np.float32, copy=True, order="C")
    d, n = p.shape
    k = c.shape[1]
    labels = np.zeros(n, np.uint32)
    masses = np.zeros(k, np.uint32)
    rc = lib().ora_lloyd3_mt_f32(pp, n, d, k, c.ctypes.data, iterations, labels.ctypes.data,
                                 masses.ctypes.data, int(threads))
    assert rc == 0
    return c, labels, masses


def reduce_u32(data, cols, rows):
    a, ap = _c(data, np.uint32)
    out = np.empty(rows, np.uint32)
    lib().ora_reduce_u32(ap, cols, rows, out.ctypes.data)
    return out


def partition_matrix(src_bytes, num_dims, buffer_size):
    src = np.frombuffer(bytes(src_bytes), np.uint8)
    num_bufs = (src.size + buffer_size - 1) // buffer_size
    dst = np.zeros(num_bufs * buffer_size, np.uint8)
    rc = lib().ora_partition_matrix(src.ctypes.data, dst.ctypes.data, src.size, num_dims, buffer_size)
    return rc, dst


# -- reference-tree C Lloyd (baselines/r_stats_kmeans.c), built into _ref/ --

_R = None


def r_ref():
    """The R stats kmeans_Lloyd(_float) compiled from the reference tree, or None."""
    global _R
    if _R is None and os.path.exists(R_REF_PATH):
        R = ctypes.CDLL(R_REF_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        R.kmeans_Lloyd_float.argtypes = [ctypes.c_void_p, ip, ip, ctypes.c_void_p, ip, ctypes.c_void_p, ip,
                                         ctypes.c_void_p, ctypes.c_void_p]
        R.kmeans_Lloyd.argtypes = R.kmeans_Lloyd_float.argtypes
        R.kmeans_Lloyd_float.restype = None
        R.kmeans_Lloyd.restype = None
        _R = R
    return _R


def r_lloyd(points, centroids, maxiter, double=False):
    """R's Lloyd: returns (centroids, 0-based labels, masses, iterations run)."""
    R = r_ref()
    dt = np.float64 if double else np.float32
    p = np.ascontiguousarray(points, dtype=dt)  # [d][n] == column-major n x d
    c = np.array(centroids, dtype=dt, copy=True, order="C")  # [d][k] == column-major k x d
    d, n = p.shape
    k = c.shape[1]
    cl = np.zeros(n, np.int32)
    nc = np.zeros(k, np.int32)
    wss = np.zeros(k, dt)
    ci = ctypes.c_int
    pn, pp_, pk, pm = ci(n), ci(d), ci(k), ci(maxiter)
    fn = R.kmeans_Lloyd if double else R.kmeans_Lloyd_float
    fn(p.ctypes.data, ctypes.byref(pn), ctypes.byref(pp_), c.ctypes.data, ctypes.byref(pk),
       cl.ctypes.data, ctypes.byref(pm), nc.ctypes.data, wss.ctypes.data)
    return c, cl - 1, nc, pm.value
