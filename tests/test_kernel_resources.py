"""Every gfx950 kernel runs from registers: the compiler's resource report
gives a scratch size of zero for each (CPU; hipcc cross-compiles).

A private array indexed at run time, a helper taking one by pointer without
being inlined, or too many live values puts a kernel's per-lane state in
scratch memory, where every access is a round trip to HBM.  The report is
the assembler's ``; ScratchSize:`` line per kernel of ``hipcc -S`` with the
flags the Makefile builds with.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "making-it-rain-webgpu_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--cuda-device-only", "-S"]
SOURCES = ["api.hip", "nonbonded.hip", "bonded.hip", "pme.hip", "integrate.hip", "warmstart.hip", "minimise.hip",
           "noalloc.hip", "barostat.hip"]


def _report(src, out_dir):
    out = os.path.join(out_dir, src + ".s")
    subprocess.run([HIPCC, *FLAGS, "-o", out, os.path.join(CSRC, src)], check=True, capture_output=True)
    text = open(out).read()
    kernels = {}
    functions = set(re.findall(r"\.type\s+(_Z\w+),@function", text))
    for m in re.finditer(r"^(_Z\w+):", text, re.M):
        if m.group(1) not in functions:
            continue  # a device variable's label
        tail = text[m.end():]
        sc = re.search(r"; ScratchSize: (\d+)", tail)
        kernels[m.group(1)] = int(sc.group(1))
    return kernels


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_no_kernel_uses_scratch_memory(tmp_path):
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as pool:
        reports = list(pool.map(lambda s: _report(s, str(tmp_path)), SOURCES))
    kernels = {k: v for r in reports for k, v in r.items()}
    assert len(kernels) >= 30, sorted(kernels)
    spilled = {k: v for k, v in kernels.items() if v}
    assert not spilled, spilled
    shutil.rmtree(str(tmp_path), ignore_errors=True)
