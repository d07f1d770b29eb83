"""GPU parity of the batched (segmented) reduce / scan / sort against the
oracle, and the batching identity: a segment reduced inside a batch gives
the same bits as the same range reduced alone."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
EPS = 2.0 ** -53


def log_uniform_offsets(nseg, seed, max_log2=16, min_len=1):
    # L = clamp(floor(2^(max_log2 * u)), min_len, 2^max_log2), u from the counter generator
    from oracle import oracle as O
    u = np.array([(O.rng_raw(seed, i) >> 11) * 2.0 ** -53 for i in range(nseg)])
    lens = np.clip(np.floor(2.0 ** (max_log2 * u)), min_len, 2 ** max_log2).astype(np.int64)
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


def ragged_offsets(rng, nseg, maxlen, zero_frac=0.1):
    lens = rng.integers(0, maxlen + 1, nseg)
    lens[rng.random(nseg) < zero_frac] = 0
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


def test_segmented_reduce_f64_bound_and_identity(raja, orc):
    rng = np.random.default_rng(1)
    off = ragged_offsets(rng, 700, 70000)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    raja.fill(raja.FILL_F64_NORMAL, x, seed=5)
    plan = raja.Segments(off)
    out = raja.segmented_reduce(plan, x, raja.operators.plus, 0.0).cpu().numpy()
    h = x.cpu().numpy()
    kn, ab = orc.seg_kn_sum_f64(h, off)
    lens = np.diff(off)
    for s in range(plan.nseg):
        bound = 2 * max(1, math.ceil(math.log2(max(lens[s], 2)))) * EPS * ab[s]
        assert abs(out[s] - kn[s]) <= bound, s
    # same bits alone and in the batch
    for s in list(range(0, plan.nseg, 37)) + [int(np.argmax(lens))]:
        alone = raja.reduce_sum(x[off[s]:off[s + 1]], 0.0).item()
        assert alone == out[s], s
    out2 = raja.segmented_reduce(plan, x, raja.operators.plus, 0.0).cpu().numpy()
    assert np.array_equal(out, out2)
    plan.close()


def test_segmented_reduce_minmax_int(raja, orc):
    rng = np.random.default_rng(2)
    off = ragged_offsets(rng, 300, 5000)
    x = torch.empty(int(off[-1]), dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x, seed=6, a=-1000, b=1000)
    h = x.cpu().numpy()
    plan = raja.Segments(off)
    for op, f in ((raja.operators.plus, lambda a, i: orc.seq_sum(a, i)),
                  (raja.operators.minimum, lambda a, i: orc.seq_min(a, i)),
                  (raja.operators.maximum, lambda a, i: orc.seq_max(a, i))):
        init = {raja.operators.plus: 3, raja.operators.minimum: 500, raja.operators.maximum: -500}[op]
        out = raja.segmented_reduce(plan, x, op, init).cpu().numpy()
        exp = [f(h[off[s]:off[s + 1]], init) for s in range(plan.nseg)]
        assert np.array_equal(out, np.array(exp, dtype=np.int32))
    plan.close()


@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_segmented_scans_exact(raja, orc, dtype):
    rng = np.random.default_rng(3)
    off = ragged_offsets(rng, 500, 30000)
    n = int(off[-1])
    x32 = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x32, seed=7, a=-5, b=16)
    x = x32.to(dtype)
    h = x.cpu().numpy()
    plan = raja.Segments(off)
    out = torch.empty_like(x)
    raja.segmented_exclusive_scan(plan, x, out, raja.operators.plus, 11)
    got = out.cpu().numpy()
    for s in range(plan.nseg):
        seg = h[off[s]:off[s + 1]]
        assert np.array_equal(got[off[s]:off[s + 1]], orc.seq_exclusive(seg, orc.PLUS, 11)), s
    raja.segmented_inclusive_scan(plan, x, out, raja.operators.maximum)
    got = out.cpu().numpy()
    for s in range(0, plan.nseg, 3):
        seg = h[off[s]:off[s + 1]]
        assert np.array_equal(got[off[s]:off[s + 1]], orc.seq_inclusive(seg, orc.MAX)), s
    y = x.clone()
    raja.segmented_exclusive_scan(plan, y, y, raja.operators.plus, 0)  # in place
    ref = np.concatenate([orc.seq_exclusive(h[off[s]:off[s + 1]], orc.PLUS, 0)
                          for s in range(plan.nseg)] or [np.zeros(0, h.dtype)])
    assert np.array_equal(y.cpu().numpy(), ref)
    plan.close()


@pytest.mark.parametrize("desc", [False, True])
def test_segmented_sort_pairs_exact(raja, orc, desc):
    rng = np.random.default_rng(4)
    lens = np.concatenate([rng.integers(0, 9000, 400), [20000, 70001]])
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    k = torch.empty(n, dtype=torch.uint32, device=DEV)
    raja.fill(raja.FILL_U32_BITS, k, seed=8, a=12)  # duplicates
    v = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32_IOTA, v, a=0)
    hk = k.view(torch.int32).cpu().numpy().view(np.uint32)
    hv = v.cpu().numpy()
    plan = raja.Segments(off)
    raja.segmented_sort_pairs(plan, k, v, raja.operators.greater if desc else raja.operators.less)
    ek, ev = orc.seg_stable_sort_pairs_u32(hk, hv, off, descending=desc)
    assert np.array_equal(k.view(torch.int32).cpu().numpy().view(np.uint32), ek)
    assert np.array_equal(v.cpu().numpy(), ev)
    plan.close()


def test_segmented_config_size(raja, orc):
    """20,000 segments, lengths log-uniform in [1, 2^16]: FP64 sums within the
    bound and bit-identical to per-segment reductions; int32 scan exact."""
    off = log_uniform_offsets(20000, 6)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    raja.fill(raja.FILL_F64_NORMAL, x, seed=6)
    plan = raja.Segments(off)
    out = raja.segmented_reduce(plan, x).cpu().numpy()
    kn, ab = orc.seg_kn_sum_f64(x.cpu().numpy(), off)
    lens = np.diff(off)
    bound = 2 * np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2)))) * EPS * ab
    assert np.all(np.abs(out - kn) <= bound)
    for s in range(0, 20000, 997):
        assert raja.reduce_sum(x[off[s]:off[s + 1]], 0.0).item() == out[s]
    xi = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, xi, seed=3, a=0, b=16)
    oi = torch.empty_like(xi)
    raja.segmented_exclusive_scan(plan, xi, oi, raja.operators.plus, 0)
    assert np.array_equal(oi.cpu().numpy(), orc.seg_exclusive_i32(xi.cpu().numpy(), off, 0))
    plan.close()


@pytest.mark.parametrize("mix", ["short_mean", "long_mean"])
@pytest.mark.parametrize("dtype", [torch.int32, torch.int64])
def test_segmented_scan_short_segments(raja, orc, dtype, mix):
    """Short segments (one wave chunk or less: 1024 or 2048 4-byte, 1536 8-byte
    elements) with the chunk and tile edges, mixed with long ones.  4-byte
    plans whose mean length is below 65536 take tiles of 8192 elements
    (chunks of 1024), the others tiles of 16384 (chunks of 2048); 8-byte
    plans take tiles of 12288.  Exact against the oracle for plus and min."""
    rng = np.random.default_rng(17)
    edges = [0, 1, 2, 63, 64, 65, 1023, 1024, 1025, 1535, 1536, 1537, 2047, 2048, 2049,
             8191, 8192, 8193, 12287, 12288, 12289, 16383, 16384, 16385, 20000, 70000]
    if mix == "short_mean":
        lens = np.floor(2.0 ** (13 * rng.random(3000))).astype(np.int64)
        lens[:len(edges)] = edges
    else:
        lens = np.array(edges + [300000] * 12, dtype=np.int64)
    rng.shuffle(lens)
    assert (lens.mean() < 65536) == (mix == "short_mean")
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    x32 = torch.empty(n, dtype=torch.int32, device=DEV)
    raja.fill(raja.FILL_I32, x32, seed=9, a=-7, b=16)
    x = x32.to(dtype)
    h = x.cpu().numpy()
    plan = raja.Segments(off)
    out = torch.empty_like(x)
    raja.segmented_exclusive_scan(plan, x, out, raja.operators.plus, 5)
    got = out.cpu().numpy()
    ref = np.concatenate([orc.seq_exclusive(h[off[s]:off[s + 1]], orc.PLUS, 5) for s in range(plan.nseg)])
    assert np.array_equal(got, ref)
    raja.segmented_inclusive_scan(plan, x, out, raja.operators.minimum)
    got = out.cpu().numpy()
    ref = np.concatenate([orc.seq_inclusive(h[off[s]:off[s + 1]], orc.MIN) for s in range(plan.nseg)])
    assert np.array_equal(got, ref)
    plan.close()


def test_segmented_scan_f64_matches_span_scans(raja):
    """An FP64 segment scanned in a batch is bit-identical to the same range
    scanned alone, for short (one-wave) and long (tiled) segments."""
    lens = np.array([1, 5, 700, 1535, 1536, 1537, 3000, 12288, 30000, 2, 0, 900], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    raja.fill(raja.FILL_F64_NORMAL, x, seed=23)
    plan = raja.Segments(off)
    out = torch.empty_like(x)
    raja.segmented_exclusive_scan(plan, x, out, raja.operators.plus, 0.25)
    got = out.cpu().numpy()
    for s in range(plan.nseg):
        a, b = int(off[s]), int(off[s + 1])
        if b == a:
            continue
        one = torch.empty(b - a, dtype=torch.float64, device=DEV)
        raja.exclusive_scan(x[a:b].clone(), one, raja.operators.plus, 0.25)
        assert np.array_equal(got[a:b].view(np.int64), one.cpu().numpy().view(np.int64)), s
    plan.close()


@pytest.mark.parametrize("kdt,with_vals,desc,small", [(torch.int32, True, False, False),
                                                      (torch.float32, False, True, False),
                                                      (torch.uint32, False, False, False),
                                                      (torch.float32, True, False, False),
                                                      (torch.uint32, True, True, True)])
def test_segmented_sort_long_segments_one_pass_per_digit(raja, kdt, with_vals, desc, small):
    """Segments longer than one block sort (8193 to 150000 keys, mixed with
    short and empty ones) are sorted together by the segmented onesweep;
    each must equal numpy's stable sort of its own keys (ascending, or the
    stable descending order).  Small keys (below 2^12) make the high digits
    the same over whole waves of the histogram kernel (counted once per
    wave)."""
    rng = np.random.default_rng(12)
    lens = np.concatenate([rng.integers(8193, 150000, 40), [0, 1, 100, 8192, 8193, 16384, 16385],
                           rng.integers(0, 8192, 60)])
    rng.shuffle(lens)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    if kdt == torch.float32:
        hk = rng.standard_normal(n).astype(np.float32)
        hk[::7] = np.round(hk[::7], 1)  # duplicates
        hk[::101] = -0.0  # equal to +0.0 under operator<: input order kept
        hk[::103] = 0.0
    elif kdt == torch.int32:
        hk = rng.integers(-3000, 3000, n).astype(np.int32)
    else:
        hk = rng.integers(0, 1 << (12 if small else 32), n, dtype=np.uint64).astype(np.uint32)
    k = torch.from_numpy(hk.view(np.int32)).to(DEV).view(kdt) if kdt == torch.uint32 \
        else torch.from_numpy(hk).to(DEV)
    v = torch.arange(n, dtype=torch.int32, device=DEV) if with_vals else None
    plan = raja.Segments(off)
    print("segmented onesweep: %d keys in %d segments, %s, vals=%s, desc=%s, small=%s"
          % (n, len(lens), kdt, with_vals, desc, small), flush=True)
    raja.segmented_sort_pairs(plan, k, v, raja.operators.greater if desc else raja.operators.less)
    torch.cuda.synchronize()
    gk = (k.view(torch.int32).cpu().numpy().view(np.uint32) if kdt == torch.uint32
          else k.cpu().numpy())
    gv = v.cpu().numpy() if with_vals else None
    for s in range(len(lens)):
        a, b = int(off[s]), int(off[s + 1])
        seg = hk[a:b]
        # stable order by key (descending: stable on the negated order)
        idx = np.argsort(-seg.astype(np.float64) if desc else seg, kind="stable")
        assert np.array_equal(gk[a:b], seg[idx]), s
        if with_vals:
            assert np.array_equal(gv[a:b], (a + idx).astype(np.int32)), s
    plan.close()


def test_segmented_sort_many_tiles_matches_span_sorts(raja):
    """A segmented onesweep over more than 8192 tiles (about 1.4e8 pairs in
    600 segments), where each histogram block takes 4 consecutive tile
    records and adds to several segments' rows: every segment must equal the
    span sort of its own pairs, bit for bit."""
    rng = np.random.default_rng(77)
    lens = rng.integers(120000, 350000, 600)
    lens[:5] = [0, 1, 8193, 16384, 16385]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    tile = 16384
    assert int(sum((int(x) + tile - 1) // tile for x in lens)) >= 4 * 2048
    print("segmented onesweep, many tiles: %d pairs in %d segments" % (n, len(lens)), flush=True)
    k = torch.empty(n, dtype=torch.uint32, device=DEV)
    raja.fill(raja.FILL_U32_BITS, k, seed=78, a=32)
    v = torch.arange(n, dtype=torch.int32, device=DEV)
    k0, v0 = k.clone(), v.clone()
    plan = raja.Segments(off)
    raja.segmented_sort_pairs(plan, k, v)
    torch.cuda.synchronize()
    for s in range(len(lens)):
        a, b = int(off[s]), int(off[s + 1])
        kk, vv = k0[a:b].clone(), v0[a:b].clone()
        if b - a > 1:
            raja.sort_pairs(kk, vv)
        assert torch.equal(k[a:b].view(torch.int32), kk.view(torch.int32)), s
        assert torch.equal(v[a:b], vv), s
    plan.close()
