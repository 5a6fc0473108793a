"""The C ABI library loads and exports every entry point include/*.h declares;
host-side pieces (strategy names, configs, generator) behave as specified.
No GPU compute calls here."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import clkmeans

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(clk_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = clkmeans.load()
    declared = declared_functions()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", clkmeans.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (clk_[a-z0-9_]+)$", out, flags=re.M))
    missing = declared - exported
    assert not missing, missing
    for name in declared:
        getattr(lib, name)


def test_strategy_ids_follow_the_factories():
    S = clkmeans
    assert S.strategy_id(S.STAGE_FUSED, "cluster_merge") == 0
    assert S.strategy_id(S.STAGE_FUSED, "feature_sum") == 1
    assert S.strategy_id(S.STAGE_LABELING, "unroll_vector") == 0
    assert S.strategy_id(S.STAGE_LABELING, "mfma_bf16") == 1  # extensions
    assert S.strategy_id(S.STAGE_LABELING, "mfma_fp16") == 2
    for i, s in enumerate(["global_atomic", "part_global", "part_local", "part_private"]):
        assert S.strategy_id(S.STAGE_MASS_UPDATE, s) == i
    for i, s in enumerate(["feature_sum", "feature_sum_pardim", "cluster_merge"]):
        assert S.strategy_id(S.STAGE_CENTROID_UPDATE, s) == i
    for stage, bad in [(S.STAGE_FUSED, "unroll_vector"), (S.STAGE_LABELING, "cluster_merge"),
                       (S.STAGE_MASS_UPDATE, "feature_sum"), (S.STAGE_CENTROID_UPDATE, "part_local")]:
        with pytest.raises(ValueError) as e:
            S.strategy_id(stage, bad)
        assert bad in str(e.value)  # the reference throws invalid_argument(strategy)


def test_generator_deterministic_and_thread_independent():
    a, ca = clkmeans.generate_blobs(10001, 3, 7, sigma=1.0, box=50.0, seed=9, threads=1)
    b, cb = clkmeans.generate_blobs(10001, 3, 7, sigma=1.0, box=50.0, seed=9, threads=7)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ca, cb)
    c, _ = clkmeans.generate_blobs(10001, 3, 7, seed=10)
    assert not np.array_equal(a, c)


def test_generator_statistics():
    n, d, k = 200000, 4, 10
    pts, cen = clkmeans.generate_blobs(n, d, k, sigma=1.0, box=50.0, seed=0)
    assert np.all(np.abs(cen) <= 50.0)
    blob = np.arange(n) % k
    resid = pts - cen[:, blob]
    assert abs(resid.mean()) < 0.01
    assert abs(resid.std() - 1.0) < 0.01
    # equal cluster sizes, interleaved: point p belongs to blob p mod k
    assert np.ptp(np.bincount(blob, minlength=k)) <= 1
