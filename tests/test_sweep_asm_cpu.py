"""The sweep's LDS DMA is issued from inline assembly that sets M0 (the DMA's
LDS address).  That is safe only while the compiler itself never uses M0 in
the sweep kernels: this test compiles dense_sweep.hip for gfx950 with the
library's flags and checks that every M0 reference in the device assembly is
one of those inline-assembly M0 writes.  It also checks that no sweep or
field-GEMM kernel uses scratch memory: scratch accesses are vector-memory ops
the sweep's explicit vmcnt waits do not count, which makes those waits wait
for them, and in either kernel each is a round trip to memory in the inner
loop."""
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


def _device_asm(tmp_path_factory, src):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    from sqaod_amd import build
    out = tmp_path_factory.mktemp("asm") / (src + ".s")
    cmd = [hipcc, *build.COMMON_FLAGS, f"-I{build.INCLUDE}", "-x", "hip", "--offload-device-only", "-S",
           os.path.join(build.CSRC, src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return out.read_text()


@pytest.fixture(scope="module")
def sweep_asm(tmp_path_factory):
    return _device_asm(tmp_path_factory, "dense_sweep.hip")


@pytest.mark.parametrize("src, pattern", [("dense_sweep.hip", "denseSweep"),
                                          ("dense_field_mfma.hip", "fieldRefresh")])
def test_kernels_use_no_scratch(tmp_path_factory, sweep_asm, src, pattern):
    """Every sweep and field-GEMM kernel's code-object metadata reports a
    private segment of 0 bytes and no VGPR spills."""
    text = sweep_asm if src == "dense_sweep.hip" else _device_asm(tmp_path_factory, src)
    kernels = re.findall(r"\.name:\s+(\S*" + pattern + r"\S*)\n(.*?)(?=\n\s+- \.|\Z)", text, flags=re.S)
    assert kernels, f"no {pattern} kernels in the metadata"
    for name, meta in kernels:
        priv = re.search(r"\.private_segment_fixed_size:\s+(\d+)", meta)
        spill = re.search(r"\.vgpr_spill_count:\s+(\d+)", meta)
        assert priv and int(priv.group(1)) == 0, f"{name} uses scratch ({priv.group(1) if priv else '?'} bytes)"
        assert spill and int(spill.group(1)) == 0, f"{name} spills VGPRs"


def test_m0_only_in_inline_dma(sweep_asm):
    """M0 is written only by the DMA helpers' inline assembly: s_mov_b32 of
    the LDS address, and s_add_u32 m0, m0, 0x100 / 0x300 / 0x400 between the
    ops of a chain."""
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
        assert in_asm and (re.fullmatch(r"s_mov_b32\s+m0,\s+s\d+", s) or
                           re.fullmatch(r"s_add_u32\s+m0,\s+m0,\s+0x[134]00", s)), f"compiler-generated M0 use: {s}"
        writes += 1
    assert writes > 0, "no inline DMA found"
