"""clk_transfer_bench: host<->device rates by buffer size, the role of the
reference's transfer_bench.cpp:96-206 (data streamed through the buffer
cache and scheduler with a trivial kernel per chunk).

--verify makes every mode check the bytes it moved (device buffer after each
H2D mode, host buffer after the D2H, host label mirror after the pipeline,
which then copies each chunk's points into its label slot); a 3 MiB buffer
over 64 MiB leaves a ragged last chunk."""
import csv
import io
import math
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "cl-kmeans_amd", "clk_transfer_bench")

COLUMNS = ["buffer_MiB", "pinned_h2d_GBps", "pageable_h2d_GBps", "pinned_d2h_GBps", "pipeline_GBps"]


def _exe():
    if not os.path.exists(EXE):
        pytest.skip("clk_transfer_bench not built")
    return EXE


def test_cli_rejects_bad_arguments():
    exe = _exe()
    assert subprocess.run([exe, "--help"], capture_output=True).returncode == 0
    assert subprocess.run([exe, "--bogus"], capture_output=True).returncode == 2
    assert subprocess.run([exe, "--size", "0"], capture_output=True).returncode == 2
    assert subprocess.run([exe, "--size"], capture_output=True).returncode == 2


@pytest.mark.gpu
def test_transfer_rates_and_data():
    sizes = [1, 3, 16, 64]
    r = subprocess.run([_exe(), "--size", "64", "--buffer-sizes", ",".join(map(str, sizes)), "--repeat", "2",
                        "--verify"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(io.StringIO(r.stdout)))
    assert [int(x["buffer_MiB"]) for x in rows] == sizes
    for row in rows:
        assert list(row) == COLUMNS
        for c in COLUMNS[1:]:
            v = float(row[c])
            # PCIe / host-memory rates: well above a stalled copy, below HBM
            assert math.isfinite(v) and 0.1 < v < 1000.0, (row, c)
