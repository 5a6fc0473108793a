"""The .conf surface (ConfigurationParser, configuration_parser.cpp:23-328):
keys, defaults, size lists, errors.  Parsing touches no GPU."""
import glob
import os

import numpy as np
import pytest

import clkmeans

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CONFS = sorted(glob.glob("/root/reference/configurations/*.conf")) + \
    (["/root/reference/test.conf"] if os.path.exists("/root/reference/test.conf") else [])
OUR_CONFS = sorted(glob.glob(os.path.join(REPO, "cl-kmeans_amd", "configurations", "*.conf")))


@pytest.mark.parametrize("path", OUR_CONFS, ids=os.path.basename)
def test_shipped_mi355x_confs_parse(path):
    c = clkmeans.conf_describe(open(path).read())
    assert c["kmeans.clusters"] == "10" and c["kmeans.iterations"] == "30"
    pipe = c["kmeans.pipeline"]
    assert pipe in ("single_stage", "three_stage", "single_stage_buffered", "three_stage_buffered")
    name = os.path.basename(path)
    assert name.startswith("amd_mi355x_%s" % pipe)
    # the variant suffix says what differs from the plain configuration
    if name.endswith("_double.conf"):
        assert (c["kmeans.types.point"], c["kmeans.types.label"], c["kmeans.types.mass"]) == \
            ("double", "uint64", "uint64")
    else:
        assert (c["kmeans.types.point"], c["kmeans.types.label"], c["kmeans.types.mass"]) == \
            ("float", "uint32", "uint32")
    if name.endswith("_bounded.conf"):
        assert c["kmeans.fused.strategy"] == "cluster_merge_bounded"


@pytest.mark.skipif(not REF_CONFS, reason="reference tree absent")
@pytest.mark.parametrize("path", REF_CONFS, ids=os.path.basename)
def test_reference_confs_parse_unchanged(path):
    """The reference's own .conf files are accepted as they are."""
    text = open(path).read()
    c = clkmeans.conf_describe(text)
    assert c["kmeans.clusters"] == "4"
    assert c["kmeans.types.point"] == "float"


def test_values_and_defaults():
    c = clkmeans.conf_describe("""
[benchmark]
runs = 3
verify = yes
[kmeans]
clusters = 7
pipeline = three_stage
iterations = 12
converge = on
[kmeans.centroid_update]
strategy = feature_sum_pardim
global_size = 8192 4
local_size = 16
local_size = 4
local_features = 4
thread_features = 2
""")
    assert c["benchmark.runs"] == "3" and c["benchmark.verify"] == "true"
    assert c["kmeans.converge"] == "true"
    assert c["kmeans.centroid_update.global_size"] == "8192,4,1"
    assert c["kmeans.centroid_update.local_size"] == "16,4,1"   # repeated key appends a dimension
    assert c["kmeans.centroid_update.thread_features"] == "2"
    assert c["kmeans.fused.strategy"] == "cluster_merge"       # default
    assert c["kmeans.types.label"] == "uint32"


@pytest.mark.parametrize("text,needle", [
    ("[kmeans]\nnot_a_key = 1\n", "unrecognised option"),
    ("[kmeans]\nclusters = ten\n", "invalid"),
    ("[kmeans]\nclusters = 1\nclusters = 2\n", "more than once"),
    ("[kmeans.fused]\nglobal_size = 1 2 3 4\n", "1 to 3"),
    ("[benchmark]\nverify = maybe\n", "invalid"),
    ("[kmeans\nclusters = 1\n", "section"),
    ("[kmeans]\nclusters\n", "key = value"),
])
def test_errors_throw_invalid_argument(text, needle):
    with pytest.raises(ValueError) as e:
        clkmeans.conf_describe(text)
    assert needle in str(e.value)


def test_kmeans_run_rejects_bad_conf_before_touching_the_gpu():
    import numpy as np
    with pytest.raises(ValueError):
        clkmeans.kmeans_run("[kmeans]\npipeline = single_stage\nbogus = 1\n",
                            np.zeros((2, 8), np.float32), np.zeros((2, 2), np.float32))


def test_type_triples_checked_before_the_gpu():
    # bench.cpp:469-503: float/uint32/uint32 or double/uint64/uint64, else
    # invalid_argument("Invalid type"); the data's type must match the .conf
    pts = np.zeros((2, 8), np.float32)
    cen = np.zeros((2, 2), np.float32)
    base = "[kmeans]\nclusters = 2\npipeline = single_stage\niterations = 1\n"
    for types, data in [("types.point = half\n", pts),
                        ("types.point = double\ntypes.label = uint32\ntypes.mass = uint32\n", pts),
                        ("types.point = double\ntypes.label = uint64\ntypes.mass = uint64\n", pts),
                        ("", pts.astype(np.float64))]:
        with pytest.raises(clkmeans.InvalidArgument) as e:
            clkmeans.kmeans_run(base + types, data, cen.astype(data.dtype))
        assert "Invalid type" in str(e.value)
