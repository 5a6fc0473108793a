"""Python access to the MI355X Lloyd-iteration library (libclkmeans.so).

The compute path is the HIP C ABI; this package only binds it (ctypes) and
provides device-memory plumbing for tests and benchmarks.
"""
from ._lib import (  # noqa: F401
    CLK_ERR_HIP, CLK_ERR_INVALID_ARGUMENT, CLK_ERR_NOMEM, CLK_ERR_SIZE, CLK_ERR_UNSUPPORTED, CLK_OK,
    EXPORTS, LIB_PATH, PKG_ROOT, STAGE_CENTROID_UPDATE, STAGE_FUSED, STAGE_LABELING, STAGE_MASS_UPDATE,
    JOB_ACTIVE, JOB_DONE, JOB_FREE, Batch, BatchConfig, ClkError, Context, InvalidArgument, KernelConfig,
    ServeRequest, SizeMismatch, check, conf_describe, generate_blobs, job_mix, kmeans_run, load,
    poisson_arrivals, serve, serve_serial, strategy_id,
)
from . import hip  # noqa: F401
