"""The sweep's LDS DMA is issued from inline assembly that sets M0 (the DMA's
LDS address).  That is safe only while the compiler itself never uses M0 in
the sweep kernels: this test compiles dense_sweep.hip for gfx950 with the
library's flags and checks that every M0 reference in the device assembly is
one of those inline-assembly M0 writes."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    return None


def test_m0_only_in_inline_dma(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    from sqaod_amd import build
    out = tmp_path / "dense_sweep.s"
    cmd = [hipcc, *build.COMMON_FLAGS, f"-I{build.INCLUDE}", "-x", "hip", "--offload-device-only", "-S",
           os.path.join(build.CSRC, "dense_sweep.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    in_asm = False
    writes = 0
    for line in out.read_text().splitlines():
        s = line.strip()
        if s.startswith(";;#ASMSTART"):
            in_asm = True
            continue
        if s.startswith(";;#ASMEND"):
            in_asm = False
            continue
        if s.startswith(";") or not re.search(r"\bm0\b", s):
            continue
        assert in_asm and re.fullmatch(r"s_mov_b32\s+m0,\s+s\d+", s), f"compiler-generated M0 use: {s}"
        writes += 1
    assert writes > 0, "no inline DMA found"
