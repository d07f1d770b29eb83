"""Pin the CPU oracle (oracle/raja_oracle.c) before trusting it.

Every group is checked against something other than itself: the golden
fixtures (closed forms of the reference tests' inline cases), numpy /
Python library routines (cumsum, argsort(kind="stable"), math.fsum,
argmin), and brute force on tiny inputs.  The `planted` tests show that a
plausible mistake in the oracle (an off-by-one exclusive scan, a
last-index-wins minloc, an unstable sort, a dropped init) fails one of those
checks.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---- generator -------------------------------------------------------------

def test_rng_published_first_output(golden):
    assert O.rng_raw(0, 0) == int(golden["rng"]["splitmix64_state0_first"])
    assert O.rng_raw(0, 0) == 0xE220A8397B1DCDAF


def test_rng_streams(golden):
    for case in golden["rng"]["streams"]:
        seed = case["seed"]
        raw = [O.rng_raw(seed, i) for i in range(16)]
        assert raw == [int(r) for r in case["raw"]]
        assert list(O.fill_u32_bits(16, seed, 32)) == case["u32_bits32"]
        assert list(O.fill_u32_bits(16, seed, 12)) == case["u32_bits12"]


def test_rng_ranges():
    x = O.fill_i32_range(100000, 3, 0, 16)
    assert x.min() == 0 and x.max() == 15
    assert abs(x.mean() - 7.5) < 0.05
    u = O.fill_f64_uniform(100000, 0, -1.0, 1.0)
    assert u.min() >= -1.0 and u.max() < 1.0
    z = O.fill_f64_normal(200000, 5)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


# ---- reducers ------------------------------------------------------------------

def test_sum_golden(golden):
    for case in golden["sum"]:
        for dt in (np.int32, np.int64, np.float64):
            x = np.array(case["data"], dtype=dt)
            for init, expect in zip(case["inits"], case["expect"]):
                assert O.seq_sum(x, init) == expect


def test_sum_closed_form():
    for n in (1, 357, 100000):
        x = np.arange(n, dtype=np.int64)
        assert O.seq_sum(x, 0) == n * (n - 1) // 2


def test_kn_sum_vs_fsum():
    rng = np.random.default_rng(1)
    for n in (10, 1000, 200000):
        x = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
        s, a = O.kn_sum_f64(x)
        exact = math.fsum(x)
        assert abs(s - exact) <= 2 * np.spacing(abs(exact)) + 1e-300
        assert a == pytest.approx(np.abs(x).sum(), rel=1e-12)
    # cancellation that a plain sum loses entirely
    x = np.array([1e100, 1.0, -1e100], dtype=np.float64)
    assert O.kn_sum_f64(x)[0] == 1.0
    assert O.seq_sum(x, 0.0) == 0.0


def test_kn_dot_vs_fsum():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 50000)
    y = rng.uniform(-1, 1, 50000)
    s, _ = O.kn_dot_f64(x, y)
    exact = math.fsum(float(a) * float(b) for a, b in zip(x, y))
    # products are rounded before fsum sees them; the compensated dot keeps
    # their errors, so allow the product-rounding error bound.
    assert abs(s - exact) <= 50000 * 2 ** -53 * np.abs(x * y).sum()


def test_minloc_maxloc_golden(golden):
    for case in golden["loc"]:
        for dt in (np.int32, np.int64, np.float64):
            x = np.array(case["data"], dtype=dt)
            assert O.seq_minloc(x) == (case["min_val"], case["min_loc"])
            y = np.array(case["data_max"], dtype=dt)
            assert O.seq_maxloc(y) == (case["max_val"], case["max_loc"])


def test_minloc_vs_numpy_first_index():
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 65537):
        x = rng.integers(0, 5, n).astype(np.float64)  # many ties
        v, l = O.seq_minloc(x)
        assert l == int(np.argmin(x)) and v == x.min()
        v, l = O.seq_maxloc(x)
        assert l == int(np.argmax(x)) and v == x.max()


def test_minloc_init_semantics():
    x = np.array([5.0, 3.0, 3.0, 9.0])
    # init smaller than all data: keeps init and its loc (-1 by default)
    assert O.seq_minloc(x, init=-1.0) == (-1.0, -1)
    # init equal to the data minimum: the initial pair wins the tie
    assert O.seq_minloc(x, init=3.0, init_loc=77) == (3.0, 77)
    # identity init: first index of the minimum
    assert O.seq_minloc(x) == (3.0, 1)
    assert O.seq_min(x, init=4.0) == 3.0 and O.seq_max(x, init=10.0) == 10.0
    # empty range: init untouched
    e = np.zeros(0)
    assert O.seq_minloc(e, init=2.0, init_loc=5) == (2.0, 5)


def test_bitwise():
    rng = np.random.default_rng(4)
    x = rng.integers(0, 2 ** 31, 1000).astype(np.int64)
    assert O.seq_bitor(x, 0) == int(np.bitwise_or.reduce(x))
    assert O.seq_bitand(x, -1) == int(np.bitwise_and.reduce(x))


# ---- scans -------------------------------------------------------------------------

_OPS = {"plus": O.PLUS, "minimum": O.MIN, "maximum": O.MAX}


def _expand(expect):
    return np.array(expect["idx"], np.int64), np.array(expect["val"], np.int64)


def test_scans_golden(golden):
    for case in golden["scan"]:
        n = case["n"]
        x = np.arange(1, n + 1, dtype=np.int64)
        op = _OPS[case["op"]]
        if case["kind"] == "inclusive":
            got = O.seq_inclusive(x, op)
        else:
            got = O.seq_exclusive(x, op, case["init"])
        idx, val = _expand(case["expect"])
        assert np.array_equal(got[idx], val), case
        assert got.size == n


def test_scans_vs_numpy():
    rng = np.random.default_rng(5)
    for dt in (np.int32, np.int64, np.float64):
        x = rng.integers(-50, 50, 5000).astype(dt)
        assert np.array_equal(O.seq_inclusive(x, O.PLUS), np.cumsum(x, dtype=dt))
        assert np.array_equal(O.seq_inclusive(x, O.MIN), np.minimum.accumulate(x))
        assert np.array_equal(O.seq_inclusive(x, O.MAX), np.maximum.accumulate(x))
        ex = O.seq_exclusive(x, O.PLUS, 7)
        assert ex[0] == 7 and np.array_equal(ex[1:], 7 + np.cumsum(x, dtype=dt)[:-1])


def test_omp_scans_equal_seq():
    x = O.fill_i32_range(1 << 20, 3, 0, 16)
    out = np.empty_like(x)
    O.omp_exclusive_i32(x, out, 0, 8)
    assert np.array_equal(out, O.seq_exclusive(x, O.PLUS, 0))
    y = O.fill_i64_range(1 << 20, 4, 0, 16)
    out64 = np.empty_like(y)
    O.omp_inclusive_i64(y, out64, 8)
    assert np.array_equal(out64, np.cumsum(y))


# ---- sorts -------------------------------------------------------------------------

def test_sorts_golden(golden):
    for case in golden["sort"]:
        if case.get("float"):
            k = np.array(case["keys"], np.float64)
            v = np.arange(case["n"], dtype=np.int32)
            ko, vo = O.seq_stable_sort_pairs(k, v)
            # the comparison sort keeps +-0.0 (equal under operator<) in input order
            assert np.array_equal(ko, np.sort(k))
            assert vo.tolist() == case["less"]
            continue
        n = case["n"]
        for dt in (np.int32, np.int64, np.float64):
            k = np.array(case["keys"], dt)
            v = np.arange(n, dtype=np.int32)
            for comp in ("less", "greater"):
                ko, vo = O.seq_stable_sort_pairs(k, v, descending=(comp == "greater"))
                order = np.array(case[comp], np.int64)
                assert np.array_equal(vo, order.astype(np.int32))
                assert np.array_equal(ko, k[order])


def test_sort_vs_numpy_stable():
    rng = np.random.default_rng(6)
    k = rng.integers(0, 1 << 12, 100000).astype(np.uint32)
    v = np.arange(k.size, dtype=np.int32)
    ko, vo = O.seq_stable_sort_pairs(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(vo, order.astype(np.int32))
    ko2, vo2 = O.omp_stable_sort_pairs(k.copy(), v.copy(), nthreads=8)
    assert np.array_equal(ko2, ko) and np.array_equal(vo2, vo)


# ---- segmented ------------------------------------------------------------------------

def test_segmented_vs_numpy():
    rng = np.random.default_rng(7)
    lens = rng.integers(0, 300, 200)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = rng.integers(-9, 9, off[-1]).astype(np.int32)
    ex = O.seg_exclusive_i32(x, off, 0)
    for s in range(lens.size):
        seg = x[off[s]:off[s + 1]]
        if seg.size:
            assert ex[off[s]] == 0
            assert np.array_equal(ex[off[s] + 1:off[s + 1]], np.cumsum(seg)[:-1])
    xf = x.astype(np.float64)
    sums = O.seg_sum_f64(xf, off)
    assert np.array_equal(sums, np.add.reduceat(xf, off[:-1]) * (lens > 0))


def test_seg_kn_sum_vs_fsum():
    rng = np.random.default_rng(8)
    lens = rng.integers(0, 2000, 300)
    lens[:3] = [0, 1, 0]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = rng.standard_normal(int(off[-1])) * np.exp(rng.uniform(-10, 10, int(off[-1])))
    s, ab = O.seg_kn_sum_f64(x, off)
    for g in range(lens.size):
        seg = x[off[g]:off[g + 1]]
        exact = math.fsum(seg)
        assert abs(s[g] - exact) <= 2 * np.spacing(abs(exact)) + 1e-300, g
        assert ab[g] == pytest.approx(np.abs(seg).sum(), rel=1e-12, abs=0.0)
    assert s[0] == 0.0 and ab[0] == 0.0 and s[1] == x[off[1]]


def test_seg_stable_sort_pairs_vs_numpy():
    rng = np.random.default_rng(9)
    lens = rng.integers(0, 700, 150)
    lens[:2] = [0, 1]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    k = rng.integers(0, 40, int(off[-1])).astype(np.uint32)  # many ties
    k[off[5]:off[6]] = rng.integers(0, 2 ** 32, int(lens[5]), dtype=np.uint64).astype(np.uint32)
    v = np.arange(k.size, dtype=np.int32)
    for desc in (False, True):
        ko, vo = O.seg_stable_sort_pairs_u32(k, v, off, descending=desc)
        for g in range(lens.size):
            a, b = off[g], off[g + 1]
            seg = k[a:b]
            # ~key reverses uint32 order and keeps ties in input order
            order = np.argsort(~seg if desc else seg, kind="stable")
            assert np.array_equal(vo[a:b], v[a:b][order]), (desc, g)
            assert np.array_equal(ko[a:b], seg[order]), (desc, g)


def test_planted_reversed_descending_sort_is_caught():
    k = np.array([3, 1, 3, 2, 1], np.uint32)
    v = np.arange(5, dtype=np.int32)
    _, vo = O.seg_stable_sort_pairs_u32(k, v, np.array([0, 5], np.int64), descending=True)
    wrong = np.argsort(k, kind="stable")[::-1]  # ties come out in reverse input order
    assert not np.array_equal(wrong.astype(np.int32), vo)
    assert list(vo) == [0, 2, 3, 1, 4]


def test_seq_dot_left_fold():
    # integer-valued data: every product and partial sum is exact
    rng = np.random.default_rng(10)
    x = rng.integers(-1000, 1000, 20000).astype(np.float64)
    y = rng.integers(-1000, 1000, 20000).astype(np.float64)
    exact = sum(int(a) * int(b) for a, b in zip(x.astype(np.int64), y.astype(np.int64)))
    assert O.seq_dot_f64(x, y, 0.0) == float(exact)
    # closed form: sum of i^2 over [0, n)
    n = 4096
    i = np.arange(n, dtype=np.float64)
    assert O.seq_dot_f64(i, i, 5.0) == 5.0 + (n - 1) * n * (2 * n - 1) // 6
    # seq rounding order: numpy's accumulate is a sequential left fold
    u = rng.uniform(-1, 1, 30000)
    w = rng.uniform(-1, 1, 30000)
    assert O.seq_dot_f64(u, w, 0.0) == np.add.accumulate(u * w)[-1]
    # a child equal to the identity leaves init untouched
    z = np.zeros(10)
    assert O.seq_dot_f64(z, z, 2.5) == 2.5


def test_planted_pairwise_dot_is_caught():
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, 30000)
    w = rng.uniform(-1, 1, 30000)
    pairwise = float(np.sum(u * w))  # numpy's pairwise summation, not a left fold
    assert pairwise != O.seq_dot_f64(u, w, 0.0)


def test_omp_multi_reduce_vs_numpy():
    rng = np.random.default_rng(12)
    for nt in (1, 3, 8):
        x = rng.integers(-50, 50, 100003).astype(np.float64)  # ties across threads
        s, mn, mx, lv, ll = O.omp_multi_reduce_f64(x, nthreads=nt)
        assert s == float(int(x.astype(np.int64).sum()))
        assert mn == x.min() and mx == x.max()
        assert (lv, ll) == (x.min(), int(np.argmin(x)))
    z = O.fill_f64_normal(1 << 18, 5)
    s = O.omp_multi_reduce_f64(z, nthreads=8)[0]
    exact = math.fsum(z)
    assert abs(s - exact) <= 2 * 18 * 2.0 ** -53 * np.abs(z).sum()
    assert O.omp_sum_f64(z, nthreads=8) == pytest.approx(exact, abs=2 * 18 * 2.0 ** -53 * np.abs(z).sum())
    assert O.omp_sum_f64(np.arange(1000, dtype=np.float64), nthreads=5) == 499500.0


def test_omp_daxpy_dot_minloc_vs_numpy():
    rng = np.random.default_rng(13)
    x = O.fill_f64_uniform(1 << 16, 0, -1.0, 1.0)
    y = O.fill_f64_uniform(1 << 16, 1, -1.0, 1.0)
    y_np = 2.5 * x + y  # numpy: a multiply then an add, two roundings
    y2 = y.copy()
    d, lv, ll = O.omp_daxpy_dot_minloc(x, y2, 2.5, nthreads=8)
    assert np.array_equal(y2, y_np)
    exact = math.fsum(x * y_np)
    assert abs(d - exact) <= 2 * 16 * 2.0 ** -53 * np.abs(x * y_np).sum()
    assert (lv, ll) == (x.min(), int(np.argmin(x)))
    # integer-valued data: the dot is exact; tied minima resolve to the first index
    xi = rng.integers(-20, 20, 50001).astype(np.float64)
    yi = rng.integers(-20, 20, 50001).astype(np.float64)
    yo = yi.copy()
    d, lv, ll = O.omp_daxpy_dot_minloc(xi, yo, 2.0, nthreads=8)
    ye = (2 * xi.astype(np.int64) + yi.astype(np.int64))
    assert np.array_equal(yo, ye.astype(np.float64))
    assert d == float(int((xi.astype(np.int64) * ye).sum()))
    assert ll == int(np.argmin(xi))


# ---- halo exchange ------------------------------------------------------------------

def test_halo_lists_golden(golden):
    g = golden["halo"]
    dims = [g["grid"]] * 3
    for kind in ("pack", "unpack"):
        lengths, lists = O.halo_lists(dims, g["halo"], kind == "unpack")
        assert list(lengths) == g[kind]["lengths"]
        starts = np.concatenate([[0], np.cumsum(lengths)])
        for l in range(26):
            assert lists[starts[l]] == g[kind]["first"][l]
            assert lists[starts[l + 1] - 1] == g[kind]["last"][l]
    assert int(np.sum(lengths)) == 798784  # 6*131072 + 12*1024 + 8*8


def test_halo_round_trip():
    dims, h, nv = [12, 10, 8], 2, 3
    plen, plist = O.halo_lists(dims, h, False)
    ulen, ulist = O.halo_lists(dims, h, True)
    assert np.array_equal(plen, ulen)
    size = (dims[0] + 2 * h) * (dims[1] + 2 * h) * (dims[2] + 2 * h)
    vars_ = [np.arange(size, dtype=np.float64) + v for v in range(nv)]
    orig = [v.copy() for v in vars_]
    O.halo_cycle(vars_, plist, ulist, plen)
    for v in range(nv):
        assert np.array_equal(vars_[v][ulist], orig[v][plist])
        untouched = np.ones(size, bool)
        untouched[ulist] = False
        assert np.array_equal(vars_[v][untouched], orig[v][untouched])
    # the omp variant moves the same bytes
    vars2 = [o.copy() for o in orig]
    O.omp_halo_cycle(vars2, plist, ulist, plen, nthreads=4)
    for v in range(nv):
        assert np.array_equal(vars2[v], vars_[v])


# ---- planted mistakes: each fails a check used above ----------------------------------

def test_planted_exclusive_off_by_one_is_caught(golden):
    case = next(c for c in golden["scan"]
                if c["n"] == 357 and c["kind"] == "exclusive" and c["op"] == "plus" and c["init"] == 13)
    x = np.arange(1, 358, dtype=np.int64)
    wrong = 13 + np.cumsum(x)  # inclusive where exclusive was meant
    idx, val = _expand(case["expect"])
    assert not np.array_equal(wrong[idx], val)


def test_planted_last_index_minloc_is_caught():
    x = np.array([2.0, 1.0, 5.0, 1.0])
    wrong_loc = int(len(x) - 1 - np.argmin(x[::-1]))  # "<=" keeps the last tie
    assert wrong_loc != O.seq_minloc(x)[1] == int(np.argmin(x))


def test_planted_unstable_sort_is_caught(golden):
    case = next(c for c in golden["sort"] if c["n"] == 1000)
    k = np.array(case["keys"], np.int64)
    wrong = np.lexsort((-np.arange(k.size), k))  # ties in reverse input order
    assert not np.array_equal(wrong, np.array(case["less"]))


def test_planted_dropped_init_is_caught(golden):
    case = golden["sum"][0]
    assert sum(case["data"]) != case["expect"][1]
