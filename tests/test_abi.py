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
    assert S.strategy_id(S.STAGE_FUSED, "cluster_merge_bounded") == 2  # extension
    assert S.strategy_id(S.STAGE_LABELING, "unroll_vector") == 0
    assert S.strategy_id(S.STAGE_LABELING, "mfma_bf16") == 1  # extensions
    assert S.strategy_id(S.STAGE_LABELING, "mfma_fp16") == 2
    assert S.strategy_id(S.STAGE_LABELING, "unroll_vector_bounded") == 3
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


def test_generator_cli_writes_the_reference_format(tmp_path):
    """clk_generate: generator.cpp's options and .bin layout (u64 d, u64 0,
    u64 n, float32 [d][n]); n = size/4/d rounded down to a multiple of k,
    then of the divisor (cluster_generator.cpp:55-62); grouped blobs by
    default; deterministic."""
    exe = os.path.join(REPO, "cl-kmeans_amd", "clk_generate")
    if not os.path.exists(exe):
        pytest.skip("clk_generate not built")
    out = [str(tmp_path / ("g%d.bin" % i)) for i in range(3)]
    args = ["--size", "1", "--features", "3", "--clusters", "4", "--radius", "0.5",
            "--domain_min", "-10", "--domain_max", "10"]
    subprocess.run([exe] + args + [out[0]], check=True)
    subprocess.run([exe] + args + [out[1]], check=True)
    subprocess.run([exe] + args + ["--interleaved", out[2]], check=True)
    head = np.fromfile(out[0], dtype=np.uint64, count=3)
    n = (1 << 20) // 4 // 3 // 4 * 4
    n -= n % 8
    assert head.tolist() == [3, 0, n]
    assert os.path.getsize(out[0]) == 24 + 4 * 3 * n
    x = np.fromfile(out[0], dtype=np.float32, offset=24).reshape(3, n)
    assert open(out[0], "rb").read() == open(out[1], "rb").read()
    assert np.all(np.abs(x) < 10 + 0.5 * 8)
    per = -(-n // 4)
    means = np.stack([x[:, b * per:(b + 1) * per].mean(axis=1) for b in range(4)], axis=1)
    spread = np.stack([x[:, b * per:(b + 1) * per].std(axis=1) for b in range(4)], axis=1)
    assert np.all(spread < 0.6) and np.all(spread > 0.4)  # each group is one blob (sigma 0.5)
    y = np.fromfile(out[2], dtype=np.float32, offset=24).reshape(3, n)
    np.testing.assert_allclose(y[:, :4].mean(axis=1) * 0, 0)
    assert np.all(np.abs(y[:, 0:4] - means) < 3)  # interleaved: point p in blob p mod 4
    csv = str(tmp_path / "g.csv")
    subprocess.run([exe, "--csv", "--size", "1", "--features", "2", "--clusters", "3", csv], check=True)
    rows = open(csv).read().splitlines()
    m = (1 << 20) // 4 // 2 // 3 * 3
    assert len(rows) == m - m % 8 and len(rows[0].split(",")) == 2
