"""Every GPU kernel of the library keeps its state in registers and LDS: no
scratch memory (private segment) and no vector registers spilled to it.  Reads the code
object metadata of the objects this repository builds (cl-kmeans_amd/build);
a spill would turn a streaming kernel into a scratch-bound one (a labeling
variant that spilled ran 25x slower)."""
import glob
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def kernel_metadata(obj, tmp):
    fat = os.path.join(tmp, os.path.basename(obj) + ".fatbin")
    co = os.path.join(tmp, os.path.basename(obj) + ".co")
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fat, obj],
                   check=True, capture_output=True)
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", "--targets=" + TARGET,
                    "--input=" + fat, "--output=" + co], check=True, capture_output=True)
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], check=True,
                           capture_output=True, text=True).stdout
    kernels = {}
    name = None
    for line in notes.splitlines():
        m = re.match(r"\s*\.name:\s+(\S+)", line)
        if m:
            name = m.group(1)
            kernels.setdefault(name, {})
            continue
        m = re.match(r"\s*\.(private_segment_fixed_size|vgpr_spill_count|sgpr_spill_count):\s+(\d+)", line)
        if m and name:
            kernels[name][m.group(1)] = int(m.group(2))
    return kernels


def test_no_kernel_uses_scratch(tmp_path):
    objs = [o for o in glob.glob(os.path.join(REPO, "cl-kmeans_amd", "build", "clk_*.o"))]
    hip_objs = [o for o in objs if os.path.exists(o.replace("/build/", "/csrc/").replace(".o", ".hip"))]
    if not hip_objs or not os.path.exists(os.path.join(LLVM, "clang-offload-bundler")):
        pytest.skip("library objects or ROCm LLVM tools not present")
    bad, seen = [], 0
    for obj in hip_objs:
        for name, md in kernel_metadata(obj, str(tmp_path)).items():
            seen += 1
            # (SGPR spills go to VGPR lanes, not to memory)
            if md.get("private_segment_fixed_size", 0) or md.get("vgpr_spill_count", 0):
                bad.append((os.path.basename(obj), name, md))
    assert seen >= 50, seen
    assert not bad, bad
    shutil.rmtree(str(tmp_path), ignore_errors=True)
