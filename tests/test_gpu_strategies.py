"""The strategy layer (C++ KmeansSingleStage / KmeansThreeStage / _buffered)
through clk_kmeans_run and the clk_bench CLI, against the oracle."""
import glob
import os
import subprocess

import numpy as np
import pytest

import clkmeans
from oracle import oracle as O
from tests.helpers import centroid_rel_err

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONF_DIR = os.path.join(REPO, "cl-kmeans_amd", "configurations")
BENCH = os.path.join(REPO, "cl-kmeans_amd", "clk_bench")


def conf(pipeline, iterations=30, k=10, **stage):
    text = "[kmeans]\nclusters = %d\npipeline = %s\niterations = %d\n" % (k, pipeline, iterations)
    for sec, kv in stage.items():
        text += "[kmeans.%s]\n" % sec + "".join("%s = %s\n" % (a, b) for a, b in kv.items())
    return text


@pytest.fixture(scope="module")
def data():
    pts, _ = clkmeans.generate_blobs(1 << 18, 4, 10, sigma=1.0, box=50.0, seed=0)
    init = O.first_x(pts, 10)
    return pts, init, O.lloyd(pts, init, 30), O.lloyd(pts, init, 30, mode=O.MODE_GRID)


CASES = [
    ("single_stage", {"fused": {"strategy": "cluster_merge"}}),
    ("single_stage", {"fused": {"strategy": "feature_sum"}}),
    ("single_stage", {"fused": {"strategy": "cluster_merge_bounded"}}),
    ("single_stage_buffered", {"fused": {"strategy": "cluster_merge_bounded"}}),
    ("three_stage", {"labeling": {"strategy": "unroll_vector"}, "mass_update": {"strategy": "part_private"},
                     "centroid_update": {"strategy": "cluster_merge"}}),
    ("three_stage", {"mass_update": {"strategy": "global_atomic"},
                     "centroid_update": {"strategy": "feature_sum_pardim", "local_features": 4}}),
    ("three_stage", {"mass_update": {"strategy": "part_local"}, "centroid_update": {"strategy": "feature_sum"}}),
    ("single_stage_buffered", {"fused": {"strategy": "cluster_merge"}}),
    ("three_stage_buffered", {"mass_update": {"strategy": "part_global"}}),
    ("three_stage", {"labeling": {"strategy": "mfma_fp16"}}),
    ("three_stage", {"labeling": {"strategy": "unroll_vector_bounded"}}),
    ("three_stage_buffered", {"labeling": {"strategy": "unroll_vector_bounded"}}),
    ("three_stage_buffered", {"labeling": {"strategy": "mfma_bf16"}, "mass_update": {"strategy": "part_local"}}),
]


@pytest.mark.parametrize("pipeline,stages", CASES, ids=["%s-%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_strategy_matches_oracle(data, pipeline, stages):
    pts, init, (rc, rl, rm), (gc, gl, gm) = data
    # 3 chunks plus a ragged tail for the buffered pipelines
    c, l, m, it, ns = clkmeans.kmeans_run(conf(pipeline, **stages), pts, init, buffer_size=4 * 4 * 100000)
    assert it == 30 and ns > 0
    np.testing.assert_array_equal(l, rl)
    np.testing.assert_array_equal(m, rm)
    ref64 = O.sums_f64(pts, rl, 10) / rm.astype(np.float64)
    assert centroid_rel_err(c, ref64, pts, rl) <= 1e-5
    # every strategy, chunked or not: the exact grid sums, so the oracle's
    # grid loop bit for bit
    np.testing.assert_array_equal(l, gl)
    np.testing.assert_array_equal(c, gc)


DOUBLE = "types.point = double\ntypes.label = uint64\ntypes.mass = uint64\n"


@pytest.mark.parametrize("pipeline,stages", [
    ("single_stage", {"fused": {"strategy": "cluster_merge"}}),
    ("three_stage", {"mass_update": {"strategy": "part_local"}, "centroid_update": {"strategy": "feature_sum"}}),
    ("single_stage_buffered", {"fused": {"strategy": "cluster_merge"}}),
    ("three_stage_buffered", {}),
])
def test_double_triple_matches_oracle(pipeline, stages):
    # the reference's Bench<double, uint64_t, uint64_t> (bench.cpp:469-481)
    n, d, k, iters = 50000, 4, 10, 12
    pts32, _ = clkmeans.generate_blobs(n, d, k, sigma=3.0, box=10.0, seed=21)
    pts = pts32.astype(np.float64) + np.random.default_rng(1).standard_normal((d, n)) * 1e-9
    init = np.ascontiguousarray(pts[:, :k])
    text = conf(pipeline, iterations=iters, k=k, **stages).replace("[kmeans]\n", "[kmeans]\n" + DOUBLE, 1)
    c, l, m, it, _ = clkmeans.kmeans_run(text, pts, init, buffer_size=8 * 4 * 15000)
    rc, rl, rm = O.lloyd_f64(pts, init, iters)
    assert it == iters and l.dtype == np.uint64 and m.dtype == np.uint64
    np.testing.assert_array_equal(l, rl)
    np.testing.assert_array_equal(m, rm)
    np.testing.assert_allclose(c, rc, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("labeling", ["mfma_fp16", "mfma_bf16"])
def test_dense_three_stage_matches_oracle(labeling):
    # d=64, k=1024: MFMA labeling with refinement and the feature-slice
    # centroid update, three iterations against the oracle's FP64-merge loop.
    n, d, k = 1 << 13, 64, 1024  # the single-threaded oracle: ~1.6e9 fma
    pts, _ = clkmeans.generate_blobs(n, d, k, sigma=0.5, box=1.0, seed=3)
    init = O.first_x(pts, k)
    rc, rl, rm = O.lloyd(pts, init, 3)
    text = conf("three_stage", iterations=3, k=k, labeling={"strategy": labeling},
                centroid_update={"strategy": "cluster_merge"})
    c, l, m, it, _ = clkmeans.kmeans_run(text, pts, init)
    assert it == 3
    np.testing.assert_array_equal(l, rl)
    np.testing.assert_array_equal(m, rm)
    with np.errstate(invalid="ignore", divide="ignore"):
        ref64 = O.sums_f64(pts, rl, k) / rm.astype(np.float64)
    assert centroid_rel_err(c, ref64, pts, rl) <= 1e-5
    np.testing.assert_array_equal(np.isnan(c).all(axis=0), rm == 0)  # empty clusters: NaN


def test_convergence_tolerance_stops_early(data):
    pts, init = data[0], data[1]
    c, l, m, it, _ = clkmeans.kmeans_run(conf("single_stage", iterations=50), pts, init, tolerance=1e-4)
    assert 1 <= it < 50
    ref_c, ref_l, ref_m = O.lloyd(pts, init, it)
    np.testing.assert_array_equal(l, ref_l)


def test_unknown_strategy_raises_invalid_argument(data):
    pts, init = data[0], data[1]
    with pytest.raises(ValueError):
        clkmeans.kmeans_run(conf("single_stage", fused={"strategy": "no_such"}), pts, init)
    with pytest.raises(ValueError):
        clkmeans.kmeans_run(conf("no_such_pipeline"), pts, init)


def write_bin(path, pts):
    d, n = pts.shape
    with open(path, "wb") as fh:
        np.array([d, 0, n], np.uint64).tofile(fh)
        np.ascontiguousarray(pts, np.float32).tofile(fh)


@pytest.mark.parametrize("conf_file", sorted(glob.glob(os.path.join(CONF_DIR, "*.conf"))), ids=os.path.basename)
def test_clk_bench_verify_and_csv(tmp_path, conf_file):
    """The bench drop-in: --verify against its naive CPU Lloyd, then --csv."""
    pts, _ = clkmeans.generate_blobs(1 << 15, 4, 10, seed=4)
    binf = str(tmp_path / "blobs.bin")
    write_bin(binf, pts)
    r = subprocess.run([BENCH, "--verbose", "--verify", "--config", conf_file, binf],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Correct labels" in r.stdout
    mse = float(r.stdout.split("Centroid MSE:")[1].split()[0])
    assert mse < 1e-8
    r = subprocess.run([BENCH, "--verbose", "--csv", str(tmp_path / "out.csv"), "--config", conf_file, binf],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(os.listdir(tmp_path))
    expm = [f for f in files if f.endswith("_out_expm.csv")]
    mnts = [f for f in files if f.endswith("_out_mnts.csv")]
    evnt = [f for f in files if f.endswith("_out_evnt.csv")]
    assert len(expm) == len(mnts) == len(evnt) == 1
    params = open(tmp_path / expm[0]).read()
    assert "NumPoints,32768" in params and "NumClusters,10" in params
    rows = open(tmp_path / mnts[0]).read().splitlines()
    assert rows[0] == "ExperimentID,Run,TypeName,Iteration,Value"
    total = [r for r in rows if ",TotalTime," in r]
    assert len(total) == 1 and int(total[0].split(",")[-1]) > 0
