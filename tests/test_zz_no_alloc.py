"""sqaod_rocm_dense_bf_search_noalloc, the dense brute-force search that
allocates nothing (a static scratch array in the library, the caller's
device buffers for the problem and the results), called through the C ABI
with device buffers of the library's HIP runtime (tests/hipmem.py) and checked against the searcher's own
entry point (sqaod_rocm_dg_bf_searcher_search): the same minimum and the same
ascending x list, bit for bit, on consecutive calls of decreasing size (the
second call must not see the first call's contents), both optimize
directions and both precisions."""

import numpy as np
import pytest

import hipmem
from problems import SEED_BASE, dense_int_W, dense_graph_8x8

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def noalloc(W, dtype, optimize=0, tile=1024):
    """The entry point on device buffers of the library's own HIP runtime,
    enqueued on the null stream."""
    from sqaod_amd.rocm import _cext
    N = W.shape[0]
    dW = hipmem.DeviceArray.of(np.ascontiguousarray(W, dtype))
    cap = min((tile + 255) // 256 * 256, 2 ** N)
    dx = hipmem.DeviceArray(cap, np.uint64)
    dE = hipmem.DeviceArray(1, dtype)
    dn = hipmem.DeviceArray(1, np.uint64)
    dx.fill_bytes(0xFF)
    dn.fill_bytes(0xFF)
    rc = _cext._lib.sqaod_rocm_dense_bf_search_noalloc(0 if dtype == np.float32 else 1, dW.ptr, N, optimize, tile,
                                                       dE.ptr, dx.ptr, dn.ptr, None)
    if rc != 0:
        raise RuntimeError(_cext._lib.sqaod_rocm_last_error().decode())
    hipmem.synchronize()
    n = int(dn.download()[0])
    assert 0 < n <= cap
    return dE.download()[0], dx.download()[:n]


def searcher(rocm, W, dtype, optimize=0, tile=1024):
    import sqaod_amd as sq
    s = rocm.dense_graph_bf_searcher(W.astype(dtype), sq.maximize if optimize else sq.minimize, dtype=dtype,
                                     tile_size=tile)
    s.search()
    return s.get_E(), np.stack(s.get_x())


def unpack(xs, N):
    shifts = np.arange(N - 1, -1, -1, dtype=np.uint64)
    return ((xs[:, None] >> shifts[None, :]) & np.uint64(1)).astype(np.int8)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_search_equals_searcher_on_consecutive_calls(rocm, dtype):
    cases = [(20, 0, 1024), (14, 0, 1024), (16, 1, 256), (8, 0, 1024)]
    for N, opt, tile in cases:
        W = dense_int_W(N, SEED_BASE + 40 + N, lo=-2, hi=2) if N != 8 else dense_graph_8x8()
        Emin, xs = noalloc(W, dtype, opt, tile)
        E_ref, x_ref = searcher(rocm, W, dtype, opt, tile)
        assert len(xs) == len(x_ref), (N, opt)
        assert np.all(E_ref == Emin), (N, opt, Emin, E_ref[0])
        assert np.array_equal(unpack(xs, N), x_ref), (N, opt)


def test_noalloc_rejects_bad_arguments_before_any_device_work(rocm):
    W = dense_int_W(4, 1)
    with pytest.raises(RuntimeError, match="N must be"):
        noalloc(np.zeros((64, 64), np.float32), np.float32)
    with pytest.raises(RuntimeError, match="tile_size"):
        noalloc(W, np.float32, tile=(1 << 26) + 1)
