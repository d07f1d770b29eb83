"""Shape tuning and its on-disk cache (raja_cdna4_autotune and the lookups
every span scan and sort makes).

Every candidate shape gives the same integer bits, so: a cold cache, a
measured one, the same cache read back from disk, and a deleted one give
identical scan and sort results, all equal to the oracle; and each candidate
shape, forced through a hand-written cache file, is exact as well.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture
def tune_dir(raja, tmp_path, monkeypatch):
    monkeypatch.setenv("RAJA_CDNA4_TUNE_DIR", str(tmp_path))
    monkeypatch.delenv("RAJA_CDNA4_AUTOTUNE", raising=False)
    raja.tune_forget()
    yield tmp_path
    raja.tune_forget()


def _scan(raja, x, exclusive):
    out = torch.empty_like(x)
    if exclusive:
        raja.exclusive_scan(x, out, raja.operators.plus, 3)
    else:
        raja.inclusive_scan(x, out, raja.operators.plus)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _sort(raja, k0, v0):
    k, v = k0.clone(), v0.clone()
    raja.sort_pairs(k, v)
    torch.cuda.synchronize()
    return k.view(torch.int32).cpu().numpy().view(np.uint32), v.cpu().numpy()


def test_cold_measured_warm_deleted_give_same_bits(raja, orc, tune_dir):
    n = (1 << 20) + 77
    x = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x, seed=5, a=-50, b=50)
    want = orc.seq_exclusive(x.cpu().numpy(), orc.PLUS, 3)
    k0 = torch.empty(n, dtype=torch.uint32, device=DEV)
    raja.fill(raja.FILL_U32_BITS, k0, seed=6, a=32)
    v0 = torch.arange(n, dtype=torch.int32, device=DEV)
    ek, ev = orc.seq_stable_sort_pairs(k0.view(torch.int32).cpu().numpy().view(np.uint32),
                                       np.arange(n, dtype=np.int32))

    path = raja.tune_cache_path()
    assert path.startswith(str(tune_dir)) and "gfx950" in path
    assert raja.tune_query(raja.TUNE_EXCLUSIVE_SCAN, torch.int32, n) == (0, False)
    cold_scan, cold_sort = _scan(raja, x, True), _sort(raja, k0, v0)
    assert np.array_equal(cold_scan, want)
    assert np.array_equal(cold_sort[0], ek) and np.array_equal(cold_sort[1], ev)

    print("autotune exclusive_scan i32 and sort_pairs u32 at n=%d" % n, flush=True)
    raja.autotune(raja.TUNE_EXCLUSIVE_SCAN, torch.int32, n)
    raja.autotune(raja.TUNE_SORT_PAIRS, torch.uint32, n)
    assert os.path.exists(path)
    sc, st = raja.tune_query(raja.TUNE_EXCLUSIVE_SCAN, torch.int32, n)
    kc, kt = raja.tune_query(raja.TUNE_SORT_PAIRS, torch.uint32, n)
    assert st and kt and sc in (0, 8, 4) and kc in (0, 16, 8)
    print("chosen: scan vectors per lane %d, sort keys per thread %d" % (sc, kc), flush=True)
    assert np.array_equal(_scan(raja, x, True), cold_scan)
    s2 = _sort(raja, k0, v0)
    assert np.array_equal(s2[0], cold_sort[0]) and np.array_equal(s2[1], cold_sort[1])

    raja.tune_forget()  # warm: read back from the file
    assert raja.tune_query(raja.TUNE_EXCLUSIVE_SCAN, torch.int32, n) == (sc, True)
    assert raja.tune_query(raja.TUNE_SORT_PAIRS, torch.uint32, n) == (kc, True)
    assert np.array_equal(_scan(raja, x, True), cold_scan)
    s3 = _sort(raja, k0, v0)
    assert np.array_equal(s3[0], cold_sort[0]) and np.array_equal(s3[1], cold_sort[1])

    os.remove(path)  # deleted: default shapes again
    raja.tune_forget()
    assert raja.tune_query(raja.TUNE_SORT_PAIRS, torch.uint32, n) == (0, False)
    assert np.array_equal(_scan(raja, x, True), cold_scan)


@pytest.mark.parametrize("dtype,choices", [(torch.int32, (8, 4)), (torch.int64, (6, 3))])
def test_every_scan_shape_exact(raja, orc, tune_dir, dtype, choices):
    n = 3_000_001
    x = torch.empty(n, dtype=dtype, device=DEV)
    raja.fill(raja.FILL_I32 if dtype == torch.int32 else raja.FILL_I64, x, seed=8, a=-9, b=16)
    h = x.cpu().numpy()
    want_ex = orc.seq_exclusive(h, orc.PLUS, 3)
    want_in = orc.seq_inclusive(h, orc.PLUS)
    path = raja.tune_cache_path()
    code = 0 if dtype == torch.int32 else 1
    b = 22  # ceil(log2(3000001))
    for c in choices:
        with open(path, "w") as f:
            f.write("# forced shapes\nexclusive_scan %d %d %d 0.0\ninclusive_scan %d %d %d 0.0\n"
                    % (code, b, c, code, b, c))
        raja.tune_forget()
        assert raja.tune_query(raja.TUNE_EXCLUSIVE_SCAN, dtype, n) == (c, True)
        print("%s scans with %d vectors per lane" % (dtype, c), flush=True)
        assert np.array_equal(_scan(raja, x, True), want_ex), c
        assert np.array_equal(_scan(raja, x, False), want_in), c


@pytest.mark.parametrize("n", [100_003, 5_000_000])
def test_every_sort_shape_exact(raja, orc, tune_dir, n):
    k0 = torch.empty(n, dtype=torch.uint32, device=DEV)
    raja.fill(raja.FILL_U32_BITS, k0, seed=9, a=32)
    v0 = torch.arange(n, dtype=torch.int32, device=DEV)
    ek, ev = orc.seq_stable_sort_pairs(k0.view(torch.int32).cpu().numpy().view(np.uint32),
                                       np.arange(n, dtype=np.int32))
    path = raja.tune_cache_path()
    b = int(np.ceil(np.log2(n)))
    for c in (16, 8):
        with open(path, "w") as f:
            f.write("sort_pairs 2 %d %d 0.0\n" % (b, c))
        raja.tune_forget()
        assert raja.tune_query(raja.TUNE_SORT_PAIRS, torch.uint32, n) == (c, True)
        print("sort_pairs n=%d with %d keys per thread" % (n, c), flush=True)
        k, v = _sort(raja, k0, v0)
        assert np.array_equal(k, ek) and np.array_equal(v, ev), c
        # descending and keys-only take the same tile
        kd = k0.clone()
        raja.sort_pairs(kd, None, raja.operators.greater)
        assert np.array_equal(kd.view(torch.int32).cpu().numpy().view(np.uint32), ek[::-1]), c


def test_autotune_on_first_call(raja, orc, tune_dir, monkeypatch):
    monkeypatch.setenv("RAJA_CDNA4_AUTOTUNE", "1")
    raja.tune_forget()
    n = 1_500_000
    x = torch.empty(n, dtype=torch.int64, device=DEV)
    raja.fill(raja.FILL_I64, x, seed=10, a=0, b=16)
    assert raja.tune_query(raja.TUNE_INCLUSIVE_SCAN, torch.int64, n)[1] is False
    out = _scan(raja, x, False)  # measures the bucket first
    assert raja.tune_query(raja.TUNE_INCLUSIVE_SCAN, torch.int64, n)[1] is True
    assert np.array_equal(out, orc.seq_inclusive(x.cpu().numpy(), orc.PLUS))
    with open(raja.tune_cache_path()) as f:
        assert "inclusive_scan 1 21 " in f.read()
