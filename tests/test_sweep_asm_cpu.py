"""The sweep's LDS DMA is issued from inline assembly that sets M0 (the DMA's
LDS address).  That is safe only while the compiler itself never uses M0 in
the sweep kernels: this test compiles dense_sweep.hip for gfx950 with the
library's flags and checks that every M0 reference in the device assembly is
one of those inline-assembly M0 writes.  It also checks that no sweep kernel
uses scratch memory: scratch accesses are vector-memory ops the sweep's
explicit vmcnt waits do not count, which makes those waits wait for them, and
each is a round trip to memory in the batch loop."""
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


@pytest.fixture(scope="module")
def sweep_asm(tmp_path_factory):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    from sqaod_amd import build
    out = tmp_path_factory.mktemp("asm") / "dense_sweep.s"
    cmd = [hipcc, *build.COMMON_FLAGS, f"-I{build.INCLUDE}", "-x", "hip", "--offload-device-only", "-S",
           os.path.join(build.CSRC, "dense_sweep.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return out.read_text()


def test_sweep_kernels_use_no_scratch(sweep_asm):
    """Every sweep kernel's code-object metadata reports a private segment of
    0 bytes and no VGPR spills."""
    kernels = re.findall(r"\.name:\s+(\S*denseSweep\S*)\n(.*?)(?=\n\s+- \.|\Z)", sweep_asm, flags=re.S)
    assert kernels, "no sweep kernels in the metadata"
    for name, meta in kernels:
        priv = re.search(r"\.private_segment_fixed_size:\s+(\d+)", meta)
        spill = re.search(r"\.vgpr_spill_count:\s+(\d+)", meta)
        assert priv and int(priv.group(1)) == 0, f"{name} uses scratch ({priv.group(1) if priv else '?'} bytes)"
        assert spill and int(spill.group(1)) == 0, f"{name} spills VGPRs"


def test_m0_only_in_inline_dma(sweep_asm):
    in_asm = False
    writes = 0
    for line in sweep_asm.splitlines():
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
