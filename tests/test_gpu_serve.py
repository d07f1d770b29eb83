"""GPU parity of the request server: every request of a mixed stream (FP64
sums, int32 exclusive scans, (uint32, int32) stable sort_pairs) checked
against the oracle, and the serving identity -- a request's result has the
same bits in a batch, alone through the span entry points, and in the
per-request mode."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
EPS = 2.0 ** -53


class Pools:
    """Inputs and outputs of a request mix, each request on its own slices."""

    def __init__(self, raja, kinds, sizes, seed, key_bits=None):
        self.kinds, self.sizes = kinds, sizes
        nsum = sum(int(n) for k, n in zip(kinds, sizes) if k == raja.REQ_SUM_F64)
        nscan = sum(int(n) for k, n in zip(kinds, sizes) if k == raja.REQ_EXCLUSIVE_SCAN_I32)
        nsort = sum(int(n) for k, n in zip(kinds, sizes) if k == raja.REQ_SORT_PAIRS_U32_I32)
        # one spare element keeps every pool non-empty
        self.x = torch.empty(nsum + 1, dtype=torch.float64, device=DEV)
        raja.fill(raja.FILL_F64_NORMAL, self.x, seed=seed)
        self.sums = torch.full((len(kinds),), float("nan"), dtype=torch.float64, device=DEV)
        self.scan_in = torch.empty(nscan + 1, dtype=torch.int32, device=DEV)
        raja.fill(raja.FILL_I32, self.scan_in, seed=seed + 1, a=0, b=16)
        self.scan_out = torch.full((nscan + 1,), -7, dtype=torch.int32, device=DEV)
        keys = torch.empty(nsort + 1, dtype=torch.int32, device=DEV)
        if key_bits is None:
            # narrow key range: many duplicates, so stability is exercised
            raja.fill(raja.FILL_I32, keys, seed=seed + 2, a=0, b=4096)
        else:
            # uint32 keys uniform over [0, 2^key_bits), stored in int32 tensors
            raja.fill(raja.FILL_U32_BITS, keys, seed=seed + 2, a=key_bits)
        self.keys = keys
        self.vals = torch.arange(nsort + 1, dtype=torch.int32, device=DEV)
        self.okeys = torch.full((nsort + 1,), -1, dtype=torch.int32, device=DEV)
        self.ovals = torch.full((nsort + 1,), -1, dtype=torch.int32, device=DEV)
        self.spans = []
        a = b = c = 0
        for k, n in zip(kinds, sizes):
            n = int(n)
            if k == raja.REQ_SUM_F64:
                self.spans.append((a, n)); a += n
            elif k == raja.REQ_EXCLUSIVE_SCAN_I32:
                self.spans.append((b, n)); b += n
            else:
                self.spans.append((c, n)); c += n

    def requests(self, raja, arrivals=None):
        reqs = []
        for i, (k, (o, n)) in enumerate(zip(self.kinds, self.spans)):
            t = 0.0 if arrivals is None else float(arrivals[i])
            if k == raja.REQ_SUM_F64:
                reqs.append(raja.request(k, n, self.x[o:], self.sums[i:], arrival_s=t))
            elif k == raja.REQ_EXCLUSIVE_SCAN_I32:
                reqs.append(raja.request(k, n, self.scan_in[o:], self.scan_out[o:], arrival_s=t))
            else:
                reqs.append(raja.request(k, n, self.keys[o:], self.okeys[o:], self.vals[o:],
                                         self.ovals[o:], arrival_s=t))
        return reqs

    def reset_outputs(self):
        self.sums.fill_(float("nan"))
        self.scan_out.fill_(-7)
        self.okeys.fill_(-1)
        self.ovals.fill_(-1)

    def snapshot(self):
        torch.cuda.synchronize()
        return (self.sums.cpu().numpy().copy(), self.scan_out.cpu().numpy().copy(),
                self.okeys.cpu().numpy().copy(), self.ovals.cpu().numpy().copy())


def check_against_oracle(raja, orc, pools, snap):
    sums, scan_out, okeys, ovals = snap
    x = pools.x.cpu().numpy()
    sin = pools.scan_in.cpu().numpy()
    keys = pools.keys.cpu().numpy().view(np.uint32)
    vals = pools.vals.cpu().numpy()
    for i, (k, (o, n)) in enumerate(zip(pools.kinds, pools.spans)):
        if k == raja.REQ_SUM_F64:
            seg = x[o:o + n]
            if n == 0:
                assert sums[i] == 0.0, i
                continue
            kn, ab = orc.kn_sum_f64(seg)
            bound = 2 * max(1, math.ceil(math.log2(max(n, 2)))) * EPS * ab
            assert abs(sums[i] - kn) <= bound, (i, n)
        elif k == raja.REQ_EXCLUSIVE_SCAN_I32:
            want = orc.seq_exclusive(sin[o:o + n], orc.PLUS, 0)
            assert np.array_equal(scan_out[o:o + n], want), (i, n)
        else:
            wk, wv = orc.seq_stable_sort_pairs(keys[o:o + n].copy(), vals[o:o + n].copy())
            assert np.array_equal(okeys[o:o + n].view(np.uint32), wk), (i, n)
            assert np.array_equal(ovals[o:o + n], wv), (i, n)


def mix(raja, nreq, seed, max_log2):
    rng = np.random.default_rng(seed)
    u = rng.random(nreq)
    kinds = np.where(u < 0.4, raja.REQ_SUM_F64,
                     np.where(u < 0.7, raja.REQ_EXCLUSIVE_SCAN_I32, raja.REQ_SORT_PAIRS_U32_I32))
    sizes = np.floor(2.0 ** (rng.random(nreq) * max_log2)).astype(np.int64)
    return [int(k) for k in kinds], sizes


def test_serve_mix_matches_oracle_and_modes_agree(raja, orc):
    kinds, sizes = mix(raja, 300, seed=11, max_log2=18)
    sizes[:6] = [0, 1, 0, 1, 8192, 8193]
    kinds[:6] = [raja.REQ_SUM_F64, raja.REQ_SORT_PAIRS_U32_I32, raja.REQ_EXCLUSIVE_SCAN_I32,
                 raja.REQ_EXCLUSIVE_SCAN_I32, raja.REQ_SORT_PAIRS_U32_I32,
                 raja.REQ_SORT_PAIRS_U32_I32]
    pools = Pools(raja, kinds, sizes, seed=21)
    print("serve mix: 300 requests, batched", flush=True)
    srv = raja.Server(raja.SERVE_BATCHED, streams=4, max_inflight=8, max_batch=64,
                      batch_limit=1 << 16, arena_bytes=1 << 30)
    res, wall = srv.run(pools.requests(raja))
    assert all(r.status == 0 for r in res)
    assert max(r.unit_size for r in res) > 1  # some requests were batched
    batched = pools.snapshot()
    check_against_oracle(raja, orc, pools, batched)
    srv.close()

    print("serve mix: per-request mode", flush=True)
    pools.reset_outputs()
    srv = raja.Server(raja.SERVE_PER_REQUEST, streams=1, max_inflight=4, max_batch=1,
                      batch_limit=0, arena_bytes=1 << 30)
    res, _ = srv.run(pools.requests(raja))
    assert all(r.status == 0 and r.unit_size == 1 for r in res)
    single = pools.snapshot()
    srv.close()
    # slots of other request kinds stay NaN in the sums array
    for a, b in zip(batched, single):
        assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f")

    # batches capped at 4 MB of algorithmic bytes: more, smaller units, same bits
    print("serve mix: batched with batch_bytes 4 MB", flush=True)
    pools.reset_outputs()
    srv = raja.Server(raja.SERVE_BATCHED, streams=4, max_inflight=8, max_batch=64,
                      batch_limit=1 << 18, arena_bytes=1 << 30, batch_bytes=4 << 20)
    res, _ = srv.run(pools.requests(raja))
    srv.close()
    assert all(r.status == 0 for r in res)
    assert max(r.unit_size for r in res) > 1  # some batches formed under the cap
    for a, b in zip(pools.snapshot(), single):
        assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f")

    # a sum served in a batch has the bits of the span reduction of its range
    for i, (k, (o, n)) in enumerate(zip(kinds, pools.spans)):
        if k == raja.REQ_SUM_F64 and i % 5 == 0:
            alone = raja.reduce_sum(pools.x[o:o + n], 0.0).item()
            assert alone == batched[0][i], (i, n)


def test_serve_admission_rejects_oversized(raja, orc):
    kinds = [raja.REQ_SUM_F64, raja.REQ_SORT_PAIRS_U32_I32, raja.REQ_EXCLUSIVE_SCAN_I32,
             raja.REQ_SORT_PAIRS_U32_I32]
    sizes = np.array([5000, 1 << 17, 3000, 100], dtype=np.int64)
    pools = Pools(raja, kinds, sizes, seed=31)
    # 256 KiB arena: the 2^17-pair sort needs about 1 MiB of alternates
    srv = raja.Server(raja.SERVE_BATCHED, streams=2, max_inflight=4, max_batch=8,
                      batch_limit=1 << 16, arena_bytes=256 << 10)
    res, _ = srv.run(pools.requests(raja))
    st = [r.status for r in res]
    assert st == [0, raja.STATUS_ADMISSION, 0, 0]
    snap = pools.snapshot()
    srv.close()
    # the rejected request left its outputs untouched; the others are right
    o, n = pools.spans[1]
    assert (snap[2][o:o + n] == -1).all()
    keep = [0, 2, 3]
    sub = Pools.__new__(Pools)
    sub.kinds = [kinds[i] for i in keep]
    sub.spans = [pools.spans[i] for i in keep]
    sub.x, sub.scan_in, sub.keys, sub.vals = pools.x, pools.scan_in, pools.keys, pools.vals
    check_against_oracle(raja, orc, sub, (snap[0][keep], snap[1], snap[2], snap[3]))


def test_serve_arrivals_and_small_arena(raja, orc):
    """Poisson arrivals honoured; an arena that holds few units at a time
    forces waiting on retirements; results stay exact."""
    kinds, sizes = mix(raja, 200, seed=41, max_log2=15)
    rng = np.random.default_rng(42)
    arrivals = np.cumsum(rng.exponential(20e-6, len(kinds)))
    pools = Pools(raja, kinds, sizes, seed=51)
    srv = raja.Server(raja.SERVE_BATCHED, streams=3, max_inflight=6, max_batch=16,
                      batch_limit=1 << 16, arena_bytes=2 << 20, honor_arrivals=True)
    res, wall = srv.run(pools.requests(raja, arrivals))
    assert all(r.status == 0 for r in res)
    for r, t in zip(res, arrivals):
        assert r.start_s >= t - 1e-9 and r.done_s >= r.start_s
    assert wall >= arrivals[-1]
    check_against_oracle(raja, orc, pools, pools.snapshot())
    srv.close()


def check_full_mix(raja, orc, pools, snap):
    """Every sum within the compensated bound, every scan exact, every sort
    by an exact size-independent check: the values (iota over the pool) are
    a permutation of the request's own, the keys follow them, are in order,
    and equal keys keep input order (the unique stable result)."""
    sums, scan_out, okeys, ovals = snap
    x = pools.x.cpu().numpy()
    sin = pools.scan_in.cpu().numpy()
    keys = pools.keys.cpu().numpy().view(np.uint32)
    for i, (k, (o, n)) in enumerate(zip(pools.kinds, pools.spans)):
        if k == raja.REQ_SUM_F64:
            kn, ab = orc.kn_sum_f64(x[o:o + n])
            bound = 2 * max(1, math.ceil(math.log2(max(n, 2)))) * EPS * ab
            assert abs(sums[i] - kn) <= bound, (i, n)
        elif k == raja.REQ_EXCLUSIVE_SCAN_I32:
            want = orc.seq_exclusive(sin[o:o + n], orc.PLUS, 0)
            assert np.array_equal(scan_out[o:o + n], want), (i, n)
        else:
            idx = ovals[o:o + n].astype(np.int64) - o
            assert idx.min() >= 0 and idx.max() < n, (i, n)
            assert np.array_equal(np.bincount(idx, minlength=n), np.ones(n, np.int64)), (i, n)
            ok = okeys[o:o + n].view(np.uint32)
            assert np.array_equal(ok, keys[o:o + n][idx]), (i, n)
            assert (ok[1:] >= ok[:-1]).all(), (i, n)
            tie = ok[1:] == ok[:-1]
            assert (idx[1:][tie] > idx[:-1][tie]).all(), (i, n)


def test_serve_full_mix_at_config_sizes(raja, orc):
    """The serving workload at its stated shape: 40% FP64 sums, 30% int32
    exclusive scans, 30% (uint32, int32) sort_pairs with keys uniform over
    [0, 2^32), n = floor(2^(10 + 14u)) up to 2^24 (one request of each kind
    at exactly 2^24), Poisson arrivals, a 64 GB arena; batched and
    per-request modes give the same bits."""
    rng = np.random.default_rng(71)
    nreq = 160
    u = rng.random(nreq)
    kinds = [int(k) for k in np.where(u < 0.4, raja.REQ_SUM_F64,
                                      np.where(u < 0.7, raja.REQ_EXCLUSIVE_SCAN_I32,
                                               raja.REQ_SORT_PAIRS_U32_I32))]
    sizes = np.floor(2.0 ** (10 + 14 * rng.random(nreq))).astype(np.int64)
    for j, k in enumerate([raja.REQ_SUM_F64, raja.REQ_EXCLUSIVE_SCAN_I32,
                           raja.REQ_SORT_PAIRS_U32_I32]):
        kinds[7 + 40 * j] = k
        sizes[7 + 40 * j] = 1 << 24
    arrivals = np.cumsum(rng.exponential(50e-6, nreq))
    print("serve full mix: %d requests, %d elements" % (nreq, int(sizes.sum())), flush=True)
    pools = Pools(raja, kinds, sizes, seed=81, key_bits=32)
    srv = raja.Server(raja.SERVE_BATCHED, streams=16, max_inflight=64, max_batch=512,
                      batch_limit=1 << 16, arena_bytes=64 << 30, honor_arrivals=True)
    res, wall = srv.run(pools.requests(raja, arrivals))
    srv.close()
    assert all(r.status == 0 for r in res)
    batched = pools.snapshot()
    raja.device_status()
    print("serve full mix: checking batched results", flush=True)
    check_full_mix(raja, orc, pools, batched)

    # every request waiting at once: units of several requests form, and
    # each result keeps its bits -- with the small ones batched, and with
    # every request batched (batch_limit 2^24: 2^24-element sums, scans and
    # sorts inside batches)
    for limit in (1 << 16, 1 << 24):
        print("serve full mix: saturation (all requests at t=0), batch_limit %d" % limit, flush=True)
        pools.reset_outputs()
        srv = raja.Server(raja.SERVE_BATCHED, streams=16, max_inflight=64, max_batch=512,
                          batch_limit=limit, arena_bytes=64 << 30)
        res, _ = srv.run(pools.requests(raja))
        srv.close()
        assert all(r.status == 0 for r in res)
        assert max(r.unit_size for r in res) > (40 if limit > 1 << 16 else 1)
        raja.device_status()
        for a, b in zip(batched, pools.snapshot()):
            assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f")

    print("serve full mix: per-request mode", flush=True)
    pools.reset_outputs()
    srv = raja.Server(raja.SERVE_PER_REQUEST, streams=1, max_inflight=8, max_batch=1,
                      batch_limit=0, arena_bytes=64 << 30, honor_arrivals=True)
    res, _ = srv.run(pools.requests(raja, arrivals))
    srv.close()
    assert all(r.status == 0 and r.unit_size == 1 for r in res)
    single = pools.snapshot()
    raja.device_status()
    for a, b in zip(batched, single):
        assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f")


def test_open_server_producer_threads(raja, orc):
    """Requests enter while the server runs: two producer threads submit a
    mixed stream in bursts (raja_cdna4_server_submit) while the main thread
    polls completions; every request completes once, bursts form batched
    units, and every result matches the oracle.  A stopped server refuses
    submissions and starts again."""
    import threading
    import time

    kinds, sizes = mix(raja, 400, seed=17, max_log2=17)
    sizes[:4] = [9000, 30000, 65536, 0]
    kinds[:4] = [raja.REQ_SORT_PAIRS_U32_I32] * 3 + [raja.REQ_SUM_F64]
    pools = Pools(raja, kinds, sizes, seed=23)
    reqs = pools.requests(raja)
    torch.cuda.synchronize()  # inputs written on torch's stream are ready
    srv = raja.Server(raja.SERVE_BATCHED, streams=4, max_inflight=16, max_batch=64,
                      batch_limit=1 << 16, arena_bytes=1 << 30)
    srv.start()
    ticket_of = {}
    lock = threading.Lock()
    errors = []

    def produce(part):
        try:
            for c in range(0, len(part), 25):
                chunk = part[c:c + 25]
                tk = srv.submit([reqs[i] for i in chunk])
                with lock:
                    for t, i in zip(tk, chunk):
                        ticket_of[t] = i
                time.sleep(0.002)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    idx = list(range(len(reqs)))
    producers = [threading.Thread(target=produce, args=(idx[p::2],)) for p in range(2)]
    print("open server: 2 producers, %d requests" % len(reqs), flush=True)
    for t in producers:
        t.start()
    done = {}
    deadline = time.time() + 120
    while len(done) < len(reqs) and time.time() < deadline:
        for tk, r in srv.poll(256, timeout_s=0.05):
            assert tk not in done
            done[tk] = (r.status, r.unit_size, r.start_s, r.done_s)
    for t in producers:
        t.join()
    srv.stop()
    assert not errors, errors
    assert len(done) == len(reqs)
    assert all(v[0] == 0 for v in done.values())
    assert all(v[3] >= v[2] for v in done.values())
    assert max(v[1] for v in done.values()) > 1  # waiting requests were batched
    assert sorted(ticket_of) == sorted(done)
    check_against_oracle(raja, orc, pools, pools.snapshot())

    with pytest.raises(raja.RajaError):
        srv.submit(reqs[0])
    # started again, the same server serves a second stream
    pools.reset_outputs()
    torch.cuda.synchronize()
    srv.start()
    tks = srv.submit(reqs)
    got = {}
    deadline = time.time() + 120
    while len(got) < len(tks) and time.time() < deadline:
        for tk, r in srv.poll(512, timeout_s=0.05):
            got[tk] = r.status
    srv.stop()
    srv.close()
    assert sorted(got) == sorted(tks) and all(v == 0 for v in got.values())
    check_against_oracle(raja, orc, pools, pools.snapshot())


def test_serve_large_sorts_skip_constant_digits(raja, orc):
    """Sorts of 2^18 + 5 pairs served alone (the out-of-place span sort)
    with keys whose high digits are constant, down to all keys equal: every
    result lands in the request's own output buffers as the stable sort, and
    the inputs are left as they were."""
    n = (1 << 18) + 5
    bits_cases = [0, 8, 16, 24, 32]
    ins, outs, reqs = [], [], []
    for j, bits in enumerate(bits_cases):
        k = torch.empty(n, dtype=torch.uint32, device=DEV)
        if bits == 0:
            raja.fill(raja.FILL_U32_CONST, k, a=12345)
        else:
            raja.fill(raja.FILL_U32_BITS, k, seed=90 + j, a=bits)
        v = torch.arange(n, dtype=torch.int32, device=DEV)
        ok = torch.full((n,), 0xFFFFFFFF, dtype=torch.int64, device=DEV).to(torch.uint32)
        ov = torch.full((n,), -1, dtype=torch.int32, device=DEV)
        ins.append((k, v))
        outs.append((ok, ov))
        reqs.append(raja.request(raja.REQ_SORT_PAIRS_U32_I32, n, k, ok, v, ov))
    print("serve sorts with constant digits: %d requests of %d pairs" % (len(reqs), n), flush=True)
    srv = raja.Server(raja.SERVE_PER_REQUEST, streams=1, max_inflight=4, max_batch=1,
                      batch_limit=0, arena_bytes=1 << 30)
    res, _ = srv.run(reqs)
    srv.close()
    assert all(r.status == 0 and r.unit_size == 1 for r in res)
    for bits, (k, v), (ok, ov) in zip(bits_cases, ins, outs):
        hk = k.view(torch.int32).cpu().numpy().view(np.uint32)
        ek, ev = orc.seq_stable_sort_pairs(hk, np.arange(n, dtype=np.int32))
        assert np.array_equal(ok.view(torch.int32).cpu().numpy().view(np.uint32), ek), bits
        assert np.array_equal(ov.cpu().numpy(), ev), bits
        # inputs untouched
        assert np.array_equal(v.cpu().numpy(), np.arange(n, dtype=np.int32)), bits
