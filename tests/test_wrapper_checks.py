"""The Python wrappers reject spans that do not match what the C entry
points will read and write (those take raw pointers and one length), before
any library call: wrong sizes, dtypes and devices raise RajaError.  Runs on
the CPU with host tensors and a stand-in plan object, so every check shown
here fires before the GPU would be touched."""
import types

import pytest
import torch

import raja_amd as R


def plan(nseg=3, total=10):
    return types.SimpleNamespace(handle=1, device=0, nseg=nseg, total=total)


def t(n, dtype=torch.float64):
    return torch.zeros(n, dtype=dtype)


def test_segmented_reduce_checks():
    with pytest.raises(R.RajaError, match="plan covers 10"):
        R.segmented_reduce(plan(), t(9))
    with pytest.raises(R.RajaError, match="needs at least 3"):
        R.segmented_reduce(plan(), t(10), out=t(2))
    with pytest.raises(R.RajaError, match="dtype"):
        R.segmented_reduce(plan(), t(10), out=t(3, torch.float32))
    with pytest.raises(R.RajaError, match="plan on cuda:0"):
        R.segmented_reduce(plan(), t(10), out=t(3))
    closed = plan()
    closed.handle = None
    with pytest.raises(R.RajaError, match="closed"):
        R.segmented_reduce(closed, t(10))


def test_segmented_scan_checks():
    for fn in (R.segmented_exclusive_scan, R.segmented_inclusive_scan):
        with pytest.raises(R.RajaError, match="plan covers 10"):
            fn(plan(), t(10, torch.int32), t(9, torch.int32))
        with pytest.raises(R.RajaError, match="plan covers 10"):
            fn(plan(), t(11, torch.int32), t(11, torch.int32))
        with pytest.raises(R.RajaError, match="same dtype"):
            fn(plan(), t(10, torch.int32), t(10, torch.int64))
        with pytest.raises(R.RajaError, match="plan on cuda:0"):
            fn(plan(), t(10, torch.int32), t(10, torch.int32))


def test_segmented_sort_checks():
    k = t(10, torch.int32)
    with pytest.raises(R.RajaError, match="plan covers 10"):
        R.segmented_sort_pairs(plan(), t(8, torch.int32), t(8, torch.int32))
    with pytest.raises(R.RajaError, match="same length"):
        R.segmented_sort_pairs(plan(), k, t(9, torch.int32))
    with pytest.raises(R.RajaError, match="4- or 8-byte"):
        R.segmented_sort_pairs(plan(), k, t(10, torch.int16))
    with pytest.raises(R.RajaError, match="less or"):
        R.segmented_sort_pairs(plan(), k, None, comparator="bogus")
    with pytest.raises(R.RajaError, match="plan on cuda:0"):
        R.segmented_sort_pairs(plan(), k, t(10, torch.int32))


def test_span_checks():
    with pytest.raises(R.RajaError, match="needs at least 1"):
        R.reduce(t(5), R.operators.plus, 0.0, out=t(0))
    with pytest.raises(R.RajaError, match="dtype"):
        R.reduce(t(5), R.operators.plus, 0.0, out=t(1, torch.int64))
    with pytest.raises(R.RajaError, match="needs at least 6"):
        R.multi_reduce_f64(t(5), out=t(5))
    with pytest.raises(R.RajaError, match="dtype"):
        R.multi_reduce_f64(t(5), out=t(6, torch.float32))
    with pytest.raises(R.RajaError, match="length and type"):
        R.exclusive_scan(t(5, torch.int32), t(4, torch.int32))
    with pytest.raises(R.RajaError, match="length and type"):
        R.inclusive_scan(t(5, torch.int32), t(5, torch.int64))
    with pytest.raises(R.RajaError, match="same length"):
        R.sort_pairs(t(5, torch.int32), t(4, torch.int32))
    with pytest.raises(R.RajaError, match="4- or 8-byte"):
        R.sort_pairs(t(5, torch.int32), t(5, torch.int16))
    with pytest.raises(R.RajaError, match="writes 8-byte"):
        R.fill(R.FILL_F64_NORMAL, t(5, torch.int32))
    with pytest.raises(R.RajaError, match="unknown fill kind"):
        R.fill(99, t(5, torch.int32))
    # sizes right: the host tensors are then refused as non-GPU spans
    with pytest.raises(R.RajaError, match="GPU tensors only"):
        R.exclusive_scan(t(5, torch.int32), t(5, torch.int32))
