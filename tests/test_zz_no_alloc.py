"""raja_cdna4_exclusive_scan_noalloc against raja_cdna4_exclusive_scan.

The allocation-free entry point keeps its lookback descriptors in a fixed
device array that every call on the GPU shares and that keeps the last
call's contents.  Two calls in a row, the second over fewer tiles than the
first, must each give the bits of the ordinary entry point (integer scans:
bit-exact), and a size beyond the array must be refused before any launch.
Both entry points are called through the C ABI.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
I32, I64 = 0, 1  # enum raja_cdna4_dtype


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)


def _scan_both(raja, L, dtype_code, op, x, init_arr):
    out_ref = torch.empty_like(x)
    out_na = torch.full_like(x, 77)
    init = init_arr.ctypes.data_as(ctypes.c_void_p)
    n = x.numel()
    st = L.raja_cdna4_exclusive_scan(dtype_code, op, x.data_ptr(), out_ref.data_ptr(), n, init,
                                     _stream())
    assert st == 0, L.raja_cdna4_last_error()
    st = L.raja_cdna4_exclusive_scan_noalloc(dtype_code, op, x.data_ptr(), out_na.data_ptr(), n,
                                             init, _stream())
    assert st == 0, L.raja_cdna4_last_error()
    torch.cuda.synchronize()
    return out_ref, out_na


@pytest.mark.parametrize("n1,n2", [(5_000_003, 70_001), (1 << 22, 32768)])
def test_noalloc_exclusive_scan_i32_two_calls(raja, orc, n1, n2):
    L = raja.lib()
    for call, n in enumerate((n1, n2)):
        print("noalloc int32 plus scan: call %d, n=%d" % (call, n), flush=True)
        x = torch.empty(n, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, x, seed=11 + call, a=-20, b=30)
        ref, na = _scan_both(raja, L, I32, raja.operators.plus, x, np.array([5], np.int32))
        assert torch.equal(ref, na), (call, n)
        # and the oracle on the same input
        want = orc.seq_exclusive(x.cpu().numpy(), orc.PLUS, 5)
        assert np.array_equal(na.cpu().numpy(), want), (call, n)


def test_noalloc_exclusive_scan_i64_max_two_calls(raja):
    L = raja.lib()
    for call, n in enumerate((3_000_017, 12_289)):
        print("noalloc int64 max scan: call %d, n=%d" % (call, n), flush=True)
        x = torch.empty(n, dtype=torch.int64, device=DEV)
        raja.fill(raja.FILL_I64, x, seed=21 + call, a=-1000000, b=1000000)
        ref, na = _scan_both(raja, L, I64, raja.operators.maximum, x,
                             np.array([-5], np.int64))
        assert torch.equal(ref, na), (call, n)
        h = x.cpu().numpy()
        want = np.maximum.accumulate(np.concatenate([[-5], h[:-1]]))
        assert np.array_equal(na.cpu().numpy(), want), (call, n)


def test_noalloc_exclusive_scan_refuses_beyond_scratch(raja):
    L = raja.lib()
    x = torch.zeros(16, dtype=torch.int64, device=DEV)
    init = np.array([0], np.int64)
    # 65537 tiles of 12288 8-byte elements: refused on the host, nothing launched
    st = L.raja_cdna4_exclusive_scan_noalloc(I64, raja.operators.plus, x.data_ptr(),
                                             x.data_ptr(), 65537 * 12288,
                                             init.ctypes.data_as(ctypes.c_void_p), _stream())
    assert st == 2  # RAJA_CDNA4_ERROR_UNSUPPORTED
    assert b"fixed scratch" in L.raja_cdna4_last_error()
    # empty range: no-op
    st = L.raja_cdna4_exclusive_scan_noalloc(I64, raja.operators.plus, x.data_ptr(),
                                             x.data_ptr(), 0,
                                             init.ctypes.data_as(ctypes.c_void_p), _stream())
    assert st == 0
