"""Shared test helpers: device round trips and the tolerance checks."""
import numpy as np

from clkmeans import hip


def dev(a):
    return hip.DeviceArray.from_host(np.ascontiguousarray(a))


def zeros(shape, dtype):
    return hip.DeviceArray.from_host(np.zeros(shape, dtype))


def centroid_rel_err(got, ref64, points, labels):
    """max |got - ref| / scale, scale = max(|ref|, mean |x| of the cluster's
    feature values): relative error that does not blow up for centroid
    coordinates near 0 (sums of mixed-sign values)."""
    d, k = ref64.shape
    absx = np.zeros((d, k))
    cnt = np.bincount(labels, minlength=k).astype(np.float64)
    for f in range(d):
        absx[f] = np.bincount(labels, weights=np.abs(points[f].astype(np.float64)), minlength=k)
    scale = np.maximum(np.abs(ref64), absx / np.maximum(cnt, 1))
    ok = cnt > 0
    err = np.abs(got.astype(np.float64) - ref64)[:, ok] / np.maximum(scale[:, ok], 1e-30)
    return float(err.max()) if err.size else 0.0
