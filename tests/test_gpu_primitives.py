"""GPU parity of the C-ABI primitives (reduce, scan, sort, fill) against the
CPU oracle on the same seeded inputs.

Bar: integer, index and sort results bit-exact; FP64 sums within
2*ceil(log2 n)*eps*sum|x| of the Kahan-Neumaier oracle (the observed error is
printed); unstable sort_pairs exact as well (the radix sort is stable).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53
DEV = "cuda:0"


def to_np(t):
    if t.dtype == torch.uint32:
        return t.view(torch.int32).cpu().numpy().view(np.uint32)
    if t.dtype == torch.uint64:
        return t.view(torch.int64).cpu().numpy().view(np.uint64)
    return t.cpu().numpy()


def to_dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).to(DEV).view(torch.uint32)
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64)).to(DEV).view(torch.uint64)
    return torch.from_numpy(a).to(DEV)


def fp_bound(n, abssum):
    return 2 * max(1, math.ceil(math.log2(max(n, 2)))) * EPS * abssum


# ---- generator ---------------------------------------------------------------------

def test_fill_matches_oracle(raja, orc):
    for n in (1, 1000, 1 << 20):
        t = torch.empty(n, dtype=torch.uint32, device=DEV)
        raja.fill(raja.FILL_U32_BITS, t, seed=1, a=32)
        assert np.array_equal(to_np(t), orc.fill_u32_bits(n, 1, 32))
        raja.fill(raja.FILL_U32_BITS, t, seed=2, a=12)
        assert np.array_equal(to_np(t), orc.fill_u32_bits(n, 2, 12))
        ti = torch.empty(n, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, ti, seed=3, a=0, b=16)
        assert np.array_equal(to_np(ti), orc.fill_i32_range(n, 3, 0, 16))
        tl = torch.empty(n, dtype=torch.int64, device=DEV)
        raja.fill(raja.FILL_I64, tl, seed=4, a=0, b=16)
        assert np.array_equal(to_np(tl), orc.fill_i64_range(n, 4, 0, 16))
        td = torch.empty(n, dtype=torch.float64, device=DEV)
        raja.fill(raja.FILL_F64_UNIFORM, td, seed=0, a=-1, b=1)
        assert np.array_equal(to_np(td), orc.fill_f64_uniform(n, 0, -1.0, 1.0))


# ---- reductions ------------------------------------------------------------------------

SIZES = [0, 1, 5, 63, 64, 4097, 8192, 8193, 100003, (1 << 20) + 7]


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64, torch.uint32, torch.float64])
def test_reduce_integer_valued_exact(raja, orc, dtype):
    for n in SIZES:
        x = torch.empty(n, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, x, seed=9, a=0, b=100)
        x = x.to(dtype) if dtype != torch.uint32 else x.view(torch.uint32)
        h = to_np(x)
        for init in (0, 2):
            got = raja.reduce_sum(x, init).item()
            assert got == orc.seq_sum(h, init), (n, init)
        assert raja.reduce_min(x).item() == orc.seq_min(h)
        assert raja.reduce_max(x).item() == orc.seq_max(h)
        assert raja.reduce_min(x, init=50).item() == orc.seq_min(h, 50)
        assert raja.reduce_max(x, init=50).item() == orc.seq_max(h, 50)


def test_reduce_bitwise(raja, orc):
    for n in (0, 1, 1000, 70001):
        x = torch.empty(n, dtype=torch.int64, device=DEV)
        raja.fill(raja.FILL_I64, x, seed=11, a=0, b=1 << 30)
        h = to_np(x)
        assert raja.reduce(x, raja.operators.bit_or, 0).item() == orc.seq_bitor(h, 0)
        assert raja.reduce(x, raja.operators.bit_and, -1).item() == orc.seq_bitand(h, -1)


def test_reduce_f64_sum_bound_and_determinism(raja, orc):
    for n in (1, 1000, 8193, (1 << 22) + 3):
        x = torch.empty(n, dtype=torch.float64, device=DEV)
        raja.fill(raja.FILL_F64_NORMAL, x, seed=5)
        h = to_np(x)
        kn, ab = orc.kn_sum_f64(h)
        r1 = raja.reduce_sum(x, 0.0).item()
        r2 = raja.reduce_sum(x, 0.0).item()
        assert r1 == r2  # bitwise identical on repeat
        err = abs(r1 - kn)
        print("reduce_sum n=%d err=%.3e bound=%.3e" % (n, err, fp_bound(n, ab)))
        assert err <= fp_bound(n, ab)


def test_reduce_unaligned_span(raja, orc):
    base = torch.empty(100001, dtype=torch.float64, device=DEV)
    raja.fill(raja.FILL_F64_IOTA, base, a=0.0)
    x = base[1:]  # 8-byte aligned, not 16
    h = to_np(x)
    assert raja.reduce_sum(x, 0.0).item() == orc.seq_sum(h, 0.0)
    v, l = raja.reduce_minloc(x)
    assert (v.item(), l.item()) == orc.seq_minloc(h)


def test_minloc_maxloc_golden(raja, golden):
    for case in golden["loc"]:
        for dt in (torch.int32, torch.int64, torch.float64, torch.float32):
            x = torch.tensor(case["data"], dtype=dt, device=DEV)
            v, l = raja.reduce_minloc(x)
            assert (v.item(), l.item()) == (case["min_val"], case["min_loc"])
            y = torch.tensor(case["data_max"], dtype=dt, device=DEV)
            v, l = raja.reduce_maxloc(y)
            assert (v.item(), l.item()) == (case["max_val"], case["max_loc"])


def test_minloc_ties_and_init(raja, orc):
    for n in (1, 64, 65, 8192 * 3 + 1, 1 << 20):
        x = torch.empty(n, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, x, seed=12, a=0, b=4)  # heavy ties
        xd = x.to(torch.float64)
        h = to_np(xd)
        v, l = raja.reduce_minloc(xd)
        assert (v.item(), l.item()) == orc.seq_minloc(h)
        v, l = raja.reduce_maxloc(xd)
        assert (v.item(), l.item()) == orc.seq_maxloc(h)
        v, l = raja.reduce_minloc(xd, init_val=-1.0, init_loc=-1)
        assert (v.item(), l.item()) == (-1.0, -1)
        v, l = raja.reduce_minloc(xd, init_val=0.0, init_loc=123)
        assert (v.item(), l.item()) == orc.seq_minloc(h, 0.0, 123)


def test_multi_reduce_f64(raja, orc):
    # up to the 2^28 FP64 N(0,1) forall of the scan + multi-reduction workload
    for n in (1, 777, (1 << 21) + 5, 1 << 28):
        print("multi_reduce_f64 n=%d" % n, flush=True)
        x = torch.empty(n, dtype=torch.float64, device=DEV)
        raja.fill(raja.FILL_F64_NORMAL, x, seed=5)
        h = to_np(x)
        r = raja.unpack_multi(raja.multi_reduce_f64(x))
        kn, ab = orc.kn_sum_f64(h)
        assert abs(r["sum"] - kn) <= fp_bound(n, ab)
        assert r["min"] == orc.seq_min(h) and r["max"] == orc.seq_max(h)
        assert (r["minloc_val"], r["minloc_loc"]) == orc.seq_minloc(h)


# ---- scans --------------------------------------------------------------------------------

_OPS = {"plus": 0, "minimum": 1, "maximum": 2}


def test_scan_golden(raja, golden):
    for case in golden["scan"]:
        n = case["n"]
        x = torch.arange(1, n + 1, dtype=torch.int64, device=DEV)
        out = torch.empty_like(x)
        op = _OPS[case["op"]]
        if case["kind"] == "inclusive":
            raja.inclusive_scan(x, out, op)
        else:
            raja.exclusive_scan(x, out, op, case["init"])
        idx = np.array(case["expect"]["idx"], np.int64)
        assert np.array_equal(to_np(out)[idx], np.array(case["expect"]["val"], np.int64)), case


# Around the 32768-element (4-byte) and 12288-element (8-byte) tiles and
# their 4096- and 1536-element wave chunks.
SCAN_SIZES = [1, 2, 100, 1535, 1537, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 12287, 12288, 12289,
              16385, 32767, 32768, 32769, 73729, 65536 * 3 + 11, (1 << 22) + 1]


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64, torch.uint32, torch.uint64])
def test_scan_integer_exact(raja, orc, dtype):
    for n in SCAN_SIZES:
        x32 = torch.empty(n, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, x32, seed=3, a=-20 if dtype in (torch.int32, torch.int64) else 0, b=16)
        if dtype == torch.uint32:
            x = x32.view(torch.uint32)
        elif dtype == torch.uint64:
            x = x32.to(torch.int64).view(torch.uint64)
        else:
            x = x32.to(dtype)
        h = to_np(x)
        for op in (0, 1, 2):
            out = torch.empty_like(x)
            raja.inclusive_scan(x, out, op)
            assert np.array_equal(to_np(out), orc.seq_inclusive(h, op)), (n, op, "inc")
            raja.exclusive_scan(x, out, op, 13 if op == 0 else None)
            exp = orc.seq_exclusive(h, op, 13 if op == 0 else None)
            assert np.array_equal(to_np(out), exp), (n, op, "exc")


def test_scan_inplace_and_unaligned(raja, orc):
    n = 300007
    x = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x, seed=21, a=0, b=16)
    h = to_np(x)
    y = x.clone()
    raja.exclusive_scan_inplace(y, 0, 5)
    assert np.array_equal(to_np(y), orc.seq_exclusive(h, 0, 5))
    y = x.clone()
    raja.inclusive_scan_inplace(y[1:], 0)  # 4-byte aligned span
    assert np.array_equal(to_np(y[1:]), orc.seq_inclusive(h[1:], 0))


def test_scan_f64_within_bound(raja, orc):
    n = (1 << 20) + 3
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    raja.fill(raja.FILL_F64_UNIFORM, x, seed=8, a=0.0, b=1.0)
    out = torch.empty_like(x)
    raja.inclusive_scan(x, out, 0)
    h = to_np(x)
    ref = np.cumsum(h.astype(np.longdouble))  # 64-bit mantissa reference
    err = np.max(np.abs(to_np(out) - ref) / np.maximum(ref, 1e-300))
    ntiles = -(-n // 4096)
    print("f64 inclusive scan max rel err %.3e" % err)
    # prefix sums chain tile aggregates: depth log2(n) inside a tile plus the tile count
    assert err <= (2 * math.ceil(math.log2(n)) + ntiles) * EPS


def test_scan_empty_is_noop(raja):
    x = torch.zeros(0, dtype=torch.int32, device=DEV)
    raja.exclusive_scan(x, x, 0, 3)
    raja.inclusive_scan(x, x, 0)


def test_scan_config_size_exact(raja, orc):
    """exclusive int32 in [0,16) and inclusive int64 at n = 2^28."""
    n = 1 << 28
    x = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x, seed=3, a=0, b=16)
    out = torch.empty_like(x)
    raja.exclusive_scan(x, out, 0, 0)
    h = to_np(x)
    ref = np.empty_like(h)
    orc.omp_exclusive_i32(h, ref, 0)
    assert np.array_equal(to_np(out), ref)
    del x, out
    y = torch.empty(n, dtype=torch.int64, device=DEV)
    raja.fill(raja.FILL_I64, y, seed=4, a=0, b=16)
    out = torch.empty_like(y)
    raja.inclusive_scan(y, out, 0)
    hy = to_np(y)
    ref = np.empty_like(hy)
    orc.omp_inclusive_i64(hy, ref)
    assert np.array_equal(to_np(out), ref)


# ---- sorts --------------------------------------------------------------------------------

def check_stable_pairs(keys_in, keys_out, vals_out, descending=False):
    """Size-independent exact check with iota values: sortedness, the
    permutation property, key/value correspondence and stability."""
    n = keys_in.size
    perm = vals_out.astype(np.int64)
    assert perm.min() >= 0 and perm.max() < n
    seen = np.zeros(n, dtype=bool)
    seen[perm] = True
    assert seen.all()  # n entries covering all n positions: a permutation
    assert np.array_equal(keys_out, keys_in[perm])
    if n > 1:
        if descending:
            assert np.all(keys_out[:-1] >= keys_out[1:])
        else:
            assert np.all(keys_out[:-1] <= keys_out[1:])
        same = keys_out[:-1] == keys_out[1:]
        assert np.all(perm[:-1][same] < perm[1:][same])


def test_sort_golden(raja, golden):
    for case in golden["sort"]:
        n = case["n"]
        if case.get("float"):
            k = torch.tensor(case["keys"], dtype=torch.float64, device=DEV)
            v = torch.arange(n, dtype=torch.int32, device=DEV)
            raja.sort_pairs(k, v)
            assert to_np(v).tolist() == case["less"]  # -0.0 and +0.0 keep input order
            continue
        for dt in (torch.int32, torch.int64, torch.float64, torch.float32):
            for comp in ("less", "greater"):
                k = torch.tensor(case["keys"], dtype=dt, device=DEV)
                v = torch.arange(n, dtype=torch.int32, device=DEV)
                raja.sort_pairs(k, v, getattr(raja.operators, comp))
                assert to_np(v).tolist() == case[comp], (n, dt, comp)
                k2 = torch.tensor(case["keys"], dtype=dt, device=DEV)
                raja.sort(k2, getattr(raja.operators, comp))
                assert np.array_equal(to_np(k2), np.array(case["keys"])[case[comp]].astype(to_np(k2).dtype))


# Around the 8192-key (64-bit keys or values) and 16384-key (32-bit) onesweep
# tiles and the 2048-key wave chunk of the latter.
SORT_SIZES = [2, 3, 255, 2047, 2049, 4095, 4096, 4097, 16383, 16384, 16385, 3 * 16384 + 2049,
              100003, (1 << 18) + 5, (1 << 21) + 17]


@pytest.mark.parametrize("bits", [32, 12, 1])
def test_sort_pairs_u32_vs_oracle(raja, orc, bits):
    for n in SORT_SIZES:
        k = torch.empty(n, dtype=torch.uint32, device=DEV)
        raja.fill(raja.FILL_U32_BITS, k, seed=1, a=bits)
        v = torch.arange(n, dtype=torch.int32, device=DEV)
        hk = to_np(k)
        raja.sort_pairs(k, v)
        ek, ev = orc.seq_stable_sort_pairs(hk, np.arange(n, dtype=np.int32))
        assert np.array_equal(to_np(k), ek), (n, bits)
        assert np.array_equal(to_np(v), ev), (n, bits)


@pytest.mark.parametrize("kdt,vdt", [(torch.int32, torch.int32), (torch.float32, torch.int64),
                                     (torch.int64, torch.int32), (torch.uint64, torch.int64),
                                     (torch.float64, torch.int64)])
def test_sort_pairs_types_descending(raja, orc, kdt, vdt):
    n = 50021
    raw = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, raw, seed=17, a=-1000, b=1000)
    if kdt == torch.uint64:
        k = raw.to(torch.int64).abs().view(torch.uint64)
    elif kdt in (torch.float32, torch.float64):
        k = raw.to(kdt) * 0.5
    else:
        k = raw.to(kdt)
    hk = to_np(k).copy()
    for desc in (False, True):
        kk = k.clone()
        v = torch.arange(n, dtype=vdt, device=DEV)
        raja.sort_pairs(kk, v, raja.operators.greater if desc else raja.operators.less)
        check_stable_pairs(hk, to_np(kk), to_np(v), desc)


@pytest.mark.parametrize("bits", [0, 8, 24])
def test_sort_pairs_skips_constant_digits(raja, orc, bits):
    """From 2^18 keys the digit passes whose digit is the same for every key
    are skipped: equal keys run none, 8-bit keys one and 24-bit keys three
    (an odd count, which in place ends with a copy back into the keys);
    every result is the stable sort."""
    for n in ((1 << 18) + 5, (1 << 20) + 3):
        print("sort skipping constant digits: n %d, %d-bit keys" % (n, bits), flush=True)
        k = torch.empty(n, dtype=torch.uint32, device=DEV)
        if bits == 0:
            raja.fill(raja.FILL_U32_CONST, k, a=77)
        else:
            raja.fill(raja.FILL_U32_BITS, k, seed=5, a=bits)
        v = torch.arange(n, dtype=torch.int32, device=DEV)
        hk = to_np(k)
        raja.sort_pairs(k, v)
        ek, ev = orc.seq_stable_sort_pairs(hk, np.arange(n, dtype=np.int32))
        assert np.array_equal(to_np(k), ek), (n, bits)
        assert np.array_equal(to_np(v), ev), (n, bits)
        raja.sort(k, raja.operators.greater)
        assert np.array_equal(to_np(k), np.sort(hk)[::-1]), (n, bits)


def test_sort_keys_only(raja, orc):
    for n in (17, 4097, 16385, 300001):
        k = torch.empty(n, dtype=torch.uint32, device=DEV)
        raja.fill(raja.FILL_U32_BITS, k, seed=31, a=32)
        hk = to_np(k)
        raja.sort(k)
        assert np.array_equal(to_np(k), np.sort(hk))
        raja.sort(k, raja.operators.greater)
        assert np.array_equal(to_np(k), np.sort(hk)[::-1])


@pytest.mark.parametrize("bits", [32, 12])
def test_sort_pairs_config_size(raja, bits):
    """N = 2^28 uint32 keys (uniform 32-bit and [0, 2^12)) with iota values."""
    n = 1 << 28
    k = torch.empty(n, dtype=torch.uint32, device=DEV)
    raja.fill(raja.FILL_U32_BITS, k, seed=1 if bits == 32 else 2, a=bits)
    hk = to_np(k)
    v = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32_IOTA, v, a=0)
    raja.sort_pairs(k, v)
    check_stable_pairs(hk, to_np(k), to_np(v))
