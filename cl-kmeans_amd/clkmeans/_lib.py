"""ctypes binding of libclkmeans.so (the C ABI declared in include/clkmeans.h)."""
import ctypes
import os

import numpy as np

from . import hip

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libclkmeans.so")

CLK_OK = 0
CLK_ERR_INVALID_ARGUMENT = -1
CLK_ERR_SIZE = -2
CLK_ERR_HIP = -3
CLK_ERR_UNSUPPORTED = -4
CLK_ERR_NOMEM = -5

STAGE_FUSED, STAGE_LABELING, STAGE_MASS_UPDATE, STAGE_CENTROID_UPDATE = 0, 1, 2, 3


class ClkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("clk error %d: %s" % (code, msg))
        self.code = code


class InvalidArgument(ClkError, ValueError):
    """The reference throws std::invalid_argument in these cases."""


class SizeMismatch(ClkError, AssertionError):
    """The reference asserts in these cases."""


class KernelConfig(ctypes.Structure):
    _fields_ = [
        ("stage", ctypes.c_int32),
        ("strategy", ctypes.c_int32),
        ("global_size", ctypes.c_uint64 * 3),
        ("local_size", ctypes.c_uint64 * 3),
        ("vector_length", ctypes.c_uint64),
        ("local_features", ctypes.c_uint64),
        ("thread_features", ctypes.c_uint64),
    ]


_lib = None

EXPORTS = [
    "clk_last_error", "clk_context_create", "clk_context_destroy", "clk_strategy_id",
    "clk_default_config", "clk_profiling_enable", "clk_profiling_set_iteration",
    "clk_profiling_clear", "clk_profiling_count", "clk_profiling_get", "clk_fused",
    "clk_labeling", "clk_mass_update", "clk_centroid_update", "clk_matrix_divide_row",
    "clk_matrix_add_u32", "clk_matrix_add_f64", "clk_reduce_vector_parcol_u32",
    "clk_reduce_vector_parcol_f64", "clk_fill_zero", "clk_init_first_x", "clk_generate_blobs",
    "clk_kmeans_run", "clk_conf_describe", "clk_lloyd_step", "clk_labeling_refined_points",
    "clk_labeling_pin_points", "clk_labeling_unpin_points",
    "clk_batch_job_bytes", "clk_batch_create", "clk_batch_destroy", "clk_batch_add", "clk_batch_step",
    "clk_batch_poll", "clk_batch_take", "clk_batch_job_buffers", "clk_batch_autotune", "clk_batch_run_serial",
    "clk_serve_run", "clk_serve_serial",
]

JOB_FREE, JOB_ACTIVE, JOB_DONE = 0, 1, 2


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("max_jobs", ctypes.c_uint32),
        ("arena_bytes", ctypes.c_uint64),
        ("tile_work", ctypes.c_uint32),
        ("max_d", ctypes.c_uint32),
        ("max_k", ctypes.c_uint32),
        ("labeling", ctypes.c_int32),
    ]


class ServeJob(ctypes.Structure):
    """clk_serve_job (clkmeans_batch.h)."""
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("d", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("points", ctypes.c_void_p),
        ("device_points", ctypes.c_int32),
        ("centroids", ctypes.c_void_p),
        ("arrival", ctypes.c_double),
        ("max_iterations", ctypes.c_uint32),
        ("tolerance", ctypes.c_double),
        ("out_centroids", ctypes.c_void_p),
        ("out_labels", ctypes.c_void_p),
        ("out_masses", ctypes.c_void_p),
        ("iterations", ctypes.c_uint32),
        ("shift", ctypes.c_double),
        ("admitted", ctypes.c_double),
        ("finished", ctypes.c_double),
    ]


class ServeStats(ctypes.Structure):
    """clk_serve_stats (clkmeans_batch.h)."""
    _fields_ = [
        ("makespan", ctypes.c_double),
        ("solves_per_s", ctypes.c_double),
        ("p50", ctypes.c_double),
        ("p99", ctypes.c_double),
        ("mean_latency", ctypes.c_double),
        ("steps", ctypes.c_uint64),
        ("max_active", ctypes.c_uint32),
        ("solo_steps", ctypes.c_uint64),
    ]


def load(path=None):
    """Load libclkmeans.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError("%s is not built; run make -C cl-kmeans_amd" % path)
    hip.lib()  # the HIP runtime first, so both resolve to one instance
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    vp, u64, i32, dp = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    cfgp = ctypes.POINTER(KernelConfig)
    sig = {
        "clk_last_error": (ctypes.c_char_p, []),
        "clk_context_create": (i32, [i32, ctypes.POINTER(vp)]),
        "clk_context_destroy": (i32, [vp]),
        "clk_strategy_id": (i32, [i32, ctypes.c_char_p]),
        "clk_default_config": (i32, [vp, i32, i32, cfgp]),
        "clk_profiling_enable": (i32, [vp, i32]),
        "clk_profiling_set_iteration": (i32, [vp, i32]),
        "clk_profiling_clear": (i32, [vp]),
        "clk_profiling_count": (ctypes.c_int64, [vp]),
        "clk_profiling_get": (i32, [vp, ctypes.c_int64, ctypes.c_char_p, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)]),
        "clk_fused": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp, vp]),
        "clk_labeling": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp]),
        "clk_lloyd_step": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "clk_lloyd_step_noalloc": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "clk_mass_update": (i32, [vp, cfgp, u64, u64, vp, vp, vp]),
        "clk_centroid_update": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp]),
        "clk_matrix_divide_row": (i32, [vp, u64, u64, vp, vp, vp, vp, vp, vp]),
        "clk_matrix_add_u32": (i32, [vp, u64, vp, vp, vp]),
        "clk_matrix_add_f64": (i32, [vp, u64, vp, vp, vp]),
        "clk_reduce_vector_parcol_u32": (i32, [vp, u64, u64, vp, vp]),
        "clk_reduce_vector_parcol_f64": (i32, [vp, u64, u64, vp, vp]),
        "clk_fill_zero": (i32, [vp, vp, u64, vp]),
        "clk_fused_f64": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp, vp]),
        "clk_lloyd_step_f64": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "clk_labeling_f64": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp]),
        "clk_mass_update_u64": (i32, [vp, cfgp, u64, u64, vp, vp, vp]),
        "clk_centroid_update_f64": (i32, [vp, cfgp, u64, u64, u64, vp, vp, vp, vp, vp]),
        "clk_matrix_divide_f64": (i32, [vp, u64, u64, vp, vp, vp, vp, vp, vp]),
        "clk_matrix_add_u64": (i32, [vp, u64, vp, vp, vp]),
        "clk_init_first_x": (i32, [vp, u64, u64, u64, vp, vp, vp]),
        "clk_sum_grid_from_points": (i32, [vp, u64, u64, vp, vp]),
        "clk_sum_grid_from_max": (i32, [vp, u64, u64, vp]),
        "clk_sum_grid_clear": (i32, [vp]),
        "clk_sum_grid_units": (i32, [vp, u64, u64, vp, vp, vp]),
        "clk_generate_blobs": (i32, [u64, u64, u64, dp, dp, u64, vp, vp, i32]),
        "clk_kmeans_run": (i32, [ctypes.c_char_p, i32, u64, u64, u64, vp, vp, vp, vp, u64, dp,
                                 ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "clk_kmeans_run_f64": (i32, [ctypes.c_char_p, i32, u64, u64, u64, vp, vp, vp, vp, u64, dp,
                                     ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "clk_conf_describe": (i32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
        "clk_labeling_refined_points": (i32, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "clk_labeling_pin_points": (i32, [vp, i32, u64, u64, vp, vp]),
        "clk_labeling_unpin_points": (i32, [vp]),
        "clk_batch_job_bytes": (u64, [ctypes.POINTER(BatchConfig), u64, ctypes.c_uint32, ctypes.c_uint32]),
        "clk_batch_create": (i32, [vp, ctypes.POINTER(BatchConfig), ctypes.POINTER(vp)]),
        "clk_batch_destroy": (i32, [vp]),
        "clk_batch_add": (i32, [vp, u64, ctypes.c_uint32, ctypes.c_uint32, vp, vp, ctypes.c_uint32, dp,
                                ctypes.POINTER(ctypes.c_int32), vp]),
        "clk_batch_step": (i32, [vp, vp]),
        "clk_batch_poll": (i32, [vp, vp, ctypes.POINTER(ctypes.c_uint32), vp]),
        "clk_batch_take": (i32, [vp, ctypes.c_int32, vp, vp, vp, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_double), vp]),
        "clk_batch_job_buffers": (i32, [vp, ctypes.c_int32, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                        ctypes.POINTER(vp)]),
        "clk_batch_autotune": (i32, [vp, ctypes.POINTER(ctypes.c_uint32), vp]),
        "clk_batch_run_serial": (i32, [vp, vp]),
        "clk_serve_run": (i32, [vp, ctypes.POINTER(ServeJob), ctypes.c_uint32, ctypes.POINTER(ServeStats), vp]),
        "clk_serve_serial": (i32, [vp, ctypes.POINTER(ServeJob), ctypes.c_uint32, ctypes.POINTER(ServeStats), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    _lib = L
    return L


def _dtype(buf):
    return np.dtype(getattr(buf, "dtype", np.float32))


def _is_f64(buf):
    return _dtype(buf) == np.float64


def check(rc):
    if rc == CLK_OK:
        return
    msg = load().clk_last_error().decode(errors="replace")
    if rc == CLK_ERR_INVALID_ARGUMENT:
        raise InvalidArgument(rc, msg)
    if rc == CLK_ERR_SIZE:
        raise SizeMismatch(rc, msg)
    raise ClkError(rc, msg)


def strategy_id(stage, name):
    """Factory strategy string -> id; InvalidArgument for unknown names."""
    rc = load().clk_strategy_id(stage, name.encode())
    if rc < 0:
        check(rc)
    return rc


def generate_blobs(n, d, k, sigma=1.0, box=50.0, seed=0, threads=None):
    """Feature-major float32 points [d][n] and blob centres [d][k] (host)."""
    threads = threads or min(16, os.cpu_count() or 1)
    pts = np.empty((d, n), np.float32)
    cen = np.empty((d, k), np.float32)
    check(load().clk_generate_blobs(n, d, k, float(sigma), float(box), int(seed),
                                    pts.ctypes.data, cen.ctypes.data, int(threads)))
    return pts, cen


def conf_describe(conf_text):
    """Parsed .conf as a dict (ConfigurationParser rules; no GPU)."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(load().clk_conf_describe(conf_text.encode(), buf, len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition("=")
        out[k] = v
    return out


def kmeans_run(conf_text, points, centroids, device=0, buffer_size=0, tolerance=0.0):
    """One clustering run of the strategy the .conf names (host arrays).
    points [d][n] float32 and centroids [d][k] float32 (initial) for the
    float/uint32/uint32 triple; float64 points select the double/uint64/uint64
    triple (the .conf must name it).  Returns (centroids, labels, masses,
    iterations_run, total_time_ns)."""
    f64 = np.asarray(points).dtype == np.float64
    pt, lt = (np.float64, np.uint64) if f64 else (np.float32, np.uint32)
    pts = np.ascontiguousarray(points, dtype=pt)
    d, n = pts.shape
    cen = np.array(centroids, dtype=pt, copy=True, order="C")
    k = cen.shape[1]
    masses = np.zeros(k, lt)
    labels = np.zeros(max(n, 1), lt)
    it = ctypes.c_uint64()
    ns = ctypes.c_uint64()
    fn = load().clk_kmeans_run_f64 if f64 else load().clk_kmeans_run
    check(fn(conf_text.encode(), device, d, n, k, pts.ctypes.data, cen.ctypes.data, masses.ctypes.data,
             labels.ctypes.data, int(buffer_size), float(tolerance), ctypes.byref(it), ctypes.byref(ns)))
    return cen, labels[:n], masses, it.value, ns.value


def _p(x):
    """Device address of a DeviceArray / int / c_void_p / None."""
    if x is None:
        return None
    if isinstance(x, hip.DeviceArray):
        return x.ptr
    if isinstance(x, ctypes.c_void_p):
        return x
    return ctypes.c_void_p(int(x))


def _s(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, hip.Stream) else stream


class Context:
    """One clk_context (owns scratch; one host thread at a time)."""

    def __init__(self, device=0):
        self.L = load()
        self.handle = ctypes.c_void_p()
        check(self.L.clk_context_create(device, ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            self.L.clk_context_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self, stage, strategy, **kw):
        """A KernelConfig from a strategy name (or id) plus overrides."""
        sid = strategy_id(stage, strategy) if isinstance(strategy, str) else int(strategy)
        cfg = KernelConfig()
        check(self.L.clk_default_config(self.handle, stage, sid, ctypes.byref(cfg)))
        for key, val in kw.items():
            if key in ("global_size", "local_size"):
                vals = list(val) if isinstance(val, (list, tuple)) else [val]
                arr = getattr(cfg, key)
                for j in range(3):
                    arr[j] = int(vals[j]) if j < len(vals) else (1 if vals and vals[0] else (0 if j == 0 else 1))
            else:
                setattr(cfg, key, int(val))
        return cfg

    # -- operators (device buffers, enqueue on `stream`) ---------------------
    def fused(self, cfg, d, n, k, points, old_centroids, sums, labels, masses, stream=None):
        """float32 points: uint32 labels/masses; float64 points: uint64 (the
        double/uint64/uint64 triple)."""
        fn = self.L.clk_fused_f64 if _is_f64(points) else self.L.clk_fused
        check(fn(self.handle, ctypes.byref(cfg) if cfg else None, d, n, k, _p(points),
                               _p(old_centroids), _p(sums), _p(labels), _p(masses), _s(stream)))

    def lloyd_step(self, cfg, d, n, k, points, old_centroids, new_centroids, sums, labels, masses,
                   max_shift=None, stream=None):
        fn = self.L.clk_lloyd_step_f64 if _is_f64(points) else self.L.clk_lloyd_step
        check(fn(self.handle, ctypes.byref(cfg) if cfg else None, d, n, k, _p(points),
                                    _p(old_centroids), _p(new_centroids), _p(sums), _p(labels), _p(masses),
                                    _p(max_shift), _s(stream)))

    def lloyd_step_noalloc(self, cfg, d, n, k, points, old_centroids, new_centroids, sums, labels, masses,
                           max_shift=None, stream=None):
        """clk_lloyd_step_noalloc: the same iteration with no device allocation."""
        check(self.L.clk_lloyd_step_noalloc(self.handle, ctypes.byref(cfg) if cfg else None, d, n, k, _p(points),
                                            _p(old_centroids), _p(new_centroids), _p(sums), _p(labels),
                                            _p(masses), _p(max_shift), _s(stream)))

    def labeling_refined_points(self):
        """Points of the last MFMA labeling call that went through the exact refinement."""
        out = ctypes.c_uint64(0)
        check(self.L.clk_labeling_refined_points(self.handle, ctypes.byref(out)))
        return out.value

    def labeling_pin_points(self, strategy, d, n, points, stream=None):
        """Build the 16-bit operand copy of device points [d][n] that later MFMA
        labelings of `strategy` ("mfma_bf16" / "mfma_fp16" or its id) read."""
        sid = strategy_id(STAGE_LABELING, strategy) if isinstance(strategy, str) else int(strategy)
        check(self.L.clk_labeling_pin_points(self.handle, sid, d, n, _p(points), _s(stream)))

    def labeling_unpin_points(self):
        check(self.L.clk_labeling_unpin_points(self.handle))

    def labeling(self, cfg, d, n, k, points, centroids, labels, stream=None):
        fn = self.L.clk_labeling_f64 if _is_f64(points) else self.L.clk_labeling
        check(fn(self.handle, ctypes.byref(cfg) if cfg else None, d, n, k, _p(points),
                                  _p(centroids), _p(labels), _s(stream)))

    def mass_update(self, cfg, n, k, labels, masses, stream=None):
        fn = self.L.clk_mass_update_u64 if _dtype(labels) == np.uint64 else self.L.clk_mass_update
        check(fn(self.handle, ctypes.byref(cfg) if cfg else None, n, k, _p(labels),
                                     _p(masses), _s(stream)))

    def centroid_update(self, cfg, d, n, k, points, sums, labels, masses=None, stream=None):
        fn = self.L.clk_centroid_update_f64 if _is_f64(points) else self.L.clk_centroid_update
        check(fn(self.handle, ctypes.byref(cfg) if cfg else None, d, n, k,
                                         _p(points), _p(sums), _p(labels), _p(masses), _s(stream)))

    def divide(self, d, k, sums, masses, centroids, old_centroids=None, max_shift=None, stream=None):
        fn = self.L.clk_matrix_divide_f64 if _is_f64(centroids) else self.L.clk_matrix_divide_row
        check(fn(self.handle, d, k, _p(sums), _p(masses), _p(centroids),
                                           _p(old_centroids), _p(max_shift), _s(stream)))

    def matrix_add(self, count, fst, snd, stream=None):
        fn = {np.dtype(np.float64): self.L.clk_matrix_add_f64,
              np.dtype(np.uint64): self.L.clk_matrix_add_u64}.get(np.dtype(fst.dtype), self.L.clk_matrix_add_u32)
        check(fn(self.handle, count, _p(fst), _p(snd), _s(stream)))

    def reduce_vector_parcol(self, num_cols, rows, data, stream=None):
        fn = (self.L.clk_reduce_vector_parcol_f64 if data.dtype == np.float64
              else self.L.clk_reduce_vector_parcol_u32)
        check(fn(self.handle, num_cols, rows, _p(data), _s(stream)))

    def fill_zero(self, buf, nbytes=None, stream=None):
        nbytes = buf.nbytes if nbytes is None else nbytes
        check(self.L.clk_fill_zero(self.handle, _p(buf), nbytes, _s(stream)))

    def first_x(self, d, n, k, points, centroids, stream=None):
        check(self.L.clk_init_first_x(self.handle, d, n, k, _p(points), _p(centroids), _s(stream)))

    # -- the sum grid (exact sums, clkmeans.h) -------------------------------
    def sum_grid_from_points(self, d, n, points, stream=None):
        """Pin the grid of device points [d][n] for this context's later ops."""
        check(self.L.clk_sum_grid_from_points(self.handle, d, n, _p(points), _s(stream)))

    def sum_grid_from_max(self, d, n, max_abs):
        m = np.ascontiguousarray(max_abs, dtype=np.float32)
        check(self.L.clk_sum_grid_from_max(self.handle, d, n, m.ctypes.data))

    def sum_grid_clear(self):
        check(self.L.clk_sum_grid_clear(self.handle))

    def sum_grid_units(self, d, n, points, stream=None):
        """The per-feature units an op would use for (d, n, points)."""
        out = np.empty(d, np.float64)
        check(self.L.clk_sum_grid_units(self.handle, d, n, _p(points), out.ctypes.data, _s(stream)))
        return out

    # -- timing records ------------------------------------------------------
    def profiling(self, enable=True):
        check(self.L.clk_profiling_enable(self.handle, 1 if enable else 0))

    def set_iteration(self, it):
        check(self.L.clk_profiling_set_iteration(self.handle, int(it)))

    def records(self):
        n = self.L.clk_profiling_count(self.handle)
        out = []
        name = ctypes.create_string_buffer(128)
        it = ctypes.c_int32()
        ms = ctypes.c_float()
        for i in range(n):
            check(self.L.clk_profiling_get(self.handle, i, name, 128, ctypes.byref(it), ctypes.byref(ms)))
            out.append((name.value.decode(), it.value, ms.value))
        return out

    def clear_records(self):
        check(self.L.clk_profiling_clear(self.handle))


class Batch:
    """Many independent k-means jobs resident on the GPU (clkmeans_batch.h)."""

    def __init__(self, ctx, max_jobs, arena_bytes, max_d, max_k, tile_work=0, labeling="unroll_vector"):
        self.L = load()
        self.ctx = ctx
        lid = strategy_id(STAGE_LABELING, labeling) if isinstance(labeling, str) else int(labeling)
        self.cfg = BatchConfig(int(max_jobs), int(arena_bytes), int(tile_work), int(max_d), int(max_k), lid)
        self.handle = ctypes.c_void_p()
        check(self.L.clk_batch_create(ctx.handle, ctypes.byref(self.cfg), ctypes.byref(self.handle)))
        self.shapes = {}

    def job_bytes(self, n, d, k):
        return self.L.clk_batch_job_bytes(ctypes.byref(self.cfg), int(n), int(d), int(k))

    def add(self, points, centroids, max_iterations, tolerance=0.0, stream=None):
        """Admit a job (host float32 [d][n] points, [d][k] initial centroids);
        returns its slot.  Raises ClkError(CLK_ERR_NOMEM) when it does not fit."""
        pts = np.ascontiguousarray(points, dtype=np.float32)
        cen = np.ascontiguousarray(centroids, dtype=np.float32)
        d, n = pts.shape
        k = cen.shape[1]
        slot = ctypes.c_int32(-1)
        check(self.L.clk_batch_add(self.handle, n, d, k, pts.ctypes.data, cen.ctypes.data, int(max_iterations),
                                   float(tolerance), ctypes.byref(slot), _s(stream)))
        self.shapes[slot.value] = (n, d, k)
        return slot.value

    def step(self, stream=None):
        check(self.L.clk_batch_step(self.handle, _s(stream)))

    def poll(self, stream=None):
        states = np.zeros(self.cfg.max_jobs, np.uint32)
        active = ctypes.c_uint32()
        check(self.L.clk_batch_poll(self.handle, states.ctypes.data, ctypes.byref(active), _s(stream)))
        return states, active.value

    def take(self, slot, stream=None):
        """(centroids [d][k], labels [n], masses [k], iterations, shift); frees the slot."""
        n, d, k = self.shapes.pop(slot)
        cen = np.empty((d, k), np.float32)
        lab = np.empty(n, np.uint32)
        mas = np.empty(k, np.uint32)
        it = ctypes.c_uint32()
        sh = ctypes.c_double()
        check(self.L.clk_batch_take(self.handle, slot, cen.ctypes.data, lab.ctypes.data, mas.ctypes.data,
                                    ctypes.byref(it), ctypes.byref(sh), _s(stream)))
        return cen, lab, mas, it.value, sh.value

    def job_buffers(self, slot):
        p, c, l = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(self.L.clk_batch_job_buffers(self.handle, slot, ctypes.byref(p), ctypes.byref(c), ctypes.byref(l)))
        return p.value, c.value, l.value

    def run_serial(self, stream=None):
        """Every active job alone through clk_lloyd_step, one after another."""
        check(self.L.clk_batch_run_serial(self.handle, _s(stream)))

    def autotune(self, stream=None):
        chosen = ctypes.c_uint32()
        check(self.L.clk_batch_autotune(self.handle, ctypes.byref(chosen), _s(stream)))
        return chosen.value

    def close(self):
        if self.handle:
            self.L.clk_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeRequest:
    """One k-means job of a serving run: host points [d][n] and initial
    centroids [d][k] (float32; pinned arrays let the copy-in overlap the GPU
    work), its arrival time in seconds and stopping rule; after the run, the
    result arrays and timings."""

    def __init__(self, points, centroids, arrival, max_iterations=50, tolerance=1e-4, want_labels=True):
        # points: host float32 [d][n], or a hip.DeviceArray (read in place)
        self.points = points
        self.centroids = centroids
        self.d, self.n = points.shape
        self.k = centroids.shape[1]
        self.on_device = not isinstance(points, np.ndarray)
        assert points.dtype == np.float32 and centroids.dtype == np.float32
        assert self.on_device or points.flags.c_contiguous
        assert centroids.flags.c_contiguous and centroids.shape[0] == self.d
        self.arrival = float(arrival)
        self.max_iterations = int(max_iterations)
        self.tolerance = float(tolerance)
        self.out_centroids = np.empty((self.d, self.k), np.float32)
        self.out_labels = np.empty(self.n, np.uint32) if want_labels else None
        self.out_masses = np.empty(self.k, np.uint32)
        self.iterations = 0
        self.shift = 0.0
        self.admitted = self.finished = 0.0

    def _struct(self):
        pts = self.points.addr if self.on_device else self.points.ctypes.data
        return ServeJob(self.n, self.d, self.k, pts, int(self.on_device), self.centroids.ctypes.data, self.arrival,
                        self.max_iterations, self.tolerance, self.out_centroids.ctypes.data,
                        self.out_labels.ctypes.data if self.out_labels is not None else None,
                        self.out_masses.ctypes.data, 0, 0.0, 0.0, 0.0)

    def _read(self, js):
        self.iterations, self.shift = js.iterations, js.shift
        self.admitted, self.finished = js.admitted, js.finished

    @property
    def latency(self):
        return self.finished - self.arrival


def _serve(fn, first, requests, stream):
    arr = (ServeJob * len(requests))(*[r._struct() for r in requests])
    st = ServeStats()
    check(fn(first, arr, len(requests), ctypes.byref(st), _s(stream)))
    for r, js in zip(requests, arr):
        r._read(js)
    return {f: getattr(st, f) for f, _ in ServeStats._fields_}


def serve(batch, requests, stream):
    """Continuous batching of the requests (arrival order) through an empty
    Batch (clk_serve_run); returns the run's statistics as a dict."""
    return _serve(batch.L.clk_serve_run, batch.handle, requests, stream)


def serve_serial(ctx, requests, stream):
    """The same requests one at a time through clk_lloyd_step (clk_serve_serial)."""
    return _serve(load().clk_serve_serial, ctx.handle, requests, stream)


def poisson_arrivals(count, rate, seed=0):
    """Arrival times (seconds) of a Poisson process of the given rate (jobs/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def job_mix(count, seed=0, log2n=(10, 18), ds=(3, 8, 16, 32), k_range=(4, 256), k_cap_div=16):
    """The batched-instances mix: N log-uniform in [2^10, 2^18], d uniform over
    {3, 8, 16, 32}, k log-uniform in [4, 256] capped at N/16; per-job seeds.
    Returns a list of (n, d, k, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(count):
        n = int(round(2.0 ** rng.uniform(*log2n)))
        d = int(rng.choice(ds))
        k = int(round(np.exp(rng.uniform(np.log(k_range[0]), np.log(k_range[1])))))
        k = max(1, min(k, n // k_cap_div))
        out.append((n, d, k, int(seed) * 1000003 + j))
    return out
