"""The device exp used by the Metropolis test equals the host C library's
expf on every one of the 2^32 float inputs (CPUDenseGraphAnnealer.cpp:262)."""
import pytest

pytestmark = pytest.mark.gpu

# Host glibc 2.35 expf checksum over all 2^32 inputs (oracle.expf_checksum),
# recorded in the build container; the GPU box runs the same image.
HOST_EXPF_CHECKSUM_FULL = 0xf3d67ff05fcc608d


def test_device_expf_exhaustive():
    from sqaod_amd.rocm import _cext
    dev = _cext.expf_checksum(0, 1 << 32)
    assert dev == HOST_EXPF_CHECKSUM_FULL


def test_device_expf_matches_host_library_here(oracle):
    """Same check against the expf of the machine the test runs on, on the
    acceptance test's domain x <= 0 (bit patterns 0x80000000..0xffffffff)."""
    from sqaod_amd.rocm import _cext
    lo, hi = 0x80000000, 1 << 32
    assert _cext.expf_checksum(lo, hi) == oracle.expf_checksum(lo, hi)
