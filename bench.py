"""Benchmark of the cdna4 hot path on one MI355X.

Headline workload: radix sort_pairs of N = 2^28 uint32 keys uniform over
[0, 2^32) with int32 iota values (the sort configuration of the reference
study).  One step = one sort_pairs call on resident inputs; the pristine keys
and values are restored by a device copy before each step, outside the
per-step event pair.  Algorithmic bytes per element = 16 per executed LSD
pass (read+write key and value) + 4 for the digit-histogram read, P = 4.

Also measured (reported under "ops"): the fused ReduceSum/Min/Max/MinLoc of
N(0,1) FP64, exclusive_scan of int32 in [0,16), inclusive_scan of int64 and
sort_pairs of keys in [0, 2^12), all at N = 2^28.

The CPU baseline is the oracle's omp_parallel_for_exec restatement of RAJA's
stable sort_pairs on a 2^26-element sample of the same distribution, timed on
this host's cores, scaled to the same GB/s definition; the reduce, scan,
forall and halo restatements are timed beside it (cpu_baseline.ops).
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline's OpenMP threads: one per core, spread (read when the
# OpenMP runtime loads, so set before torch or the oracle is imported).
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

PEAK_GBS = 8000.0
METRIC = "GB/s & %HBM peak: reduce/scan/sort_pairs 2^28; batched req/s + p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def event_time(fn, stream, reps=5, prep=None):
    import torch
    times = []
    for _ in range(reps):
        if prep is not None:
            prep()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ops", action="store_true", help="skip the secondary ops")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-serving stream")
    ap.add_argument("--no-halo", action="store_true", help="skip the WorkGroup halo exchange")
    ap.add_argument("--no-per-segment", action="store_true",
                    help="skip the one-call-per-segment comparison (tens of thousands of launches; for counter runs)")
    ap.add_argument("--no-forall", action="store_true", help="skip the forall example programs (daxpy + dot + MinLoc, Sum/Min/Max/MinLoc)")
    ap.add_argument("--no-tune", action="store_true",
                    help="take the shapes the cache holds (defaults where absent) without measuring")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run 3 plain 1 GiB device copies (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import raja_amd as R

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    n = 1 << args.log2n

    # ---- shapes: measured (or read from the on-disk cache) for the timed ops -------
    shapes = tune_shapes(R, torch, n, measure=not args.no_tune)

    # ---- headline: sort_pairs u32 uniform, iota values ------------------------------
    log("bench: sort_pairs n=2^%d setup" % args.log2n)
    keys0 = torch.empty(n, dtype=torch.uint32, device=dev)
    R.fill(R.FILL_U32_BITS, keys0, seed=1, a=32)
    vals0 = torch.empty(n, dtype=torch.int32, device=dev)
    R.fill(R.FILL_I32_IOTA, vals0, a=0)
    keys = torch.empty_like(keys0)
    vals = torch.empty_like(vals0)

    def restore():
        keys.copy_(keys0)
        vals.copy_(vals0)

    def step():
        R.sort_pairs(keys, vals)

    for _ in range(args.warmup):
        restore()
        step()
    torch.cuda.synchronize()
    R.timing_reset()
    R.timing_enable(True)
    step_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        restore()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        b.synchronize()
        step_ms.append(a.elapsed_time(b))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    R.timing_enable(False)
    passes = R.timing_query("onesweep_pass")
    hist = R.timing_query("radix_hist")
    p_exec = passes["count"] // max(1, args.steps)
    bytes_per_elem = 16 * p_exec + 4
    ms = statistics.mean(step_ms)
    gbs = n * bytes_per_elem / (ms * 1e-3) / 1e9
    pass_ms = passes["total_ms"] / max(1, passes["count"])
    pass_gbs = n * 16 / (pass_ms * 1e-3) / 1e9
    log("bench: sort_pairs %.3f ms/step (%.1f GB/s, %d passes), onesweep pass %.3f ms (%.1f GB/s)"
        % (ms, gbs, p_exec, pass_ms, pass_gbs))

    # correctness of the last step (size-independent, exact)
    hk0 = keys0.view(torch.int32).cpu().numpy().view(np.uint32)
    hk = keys.view(torch.int32).cpu().numpy().view(np.uint32)
    hv = vals.cpu().numpy().astype(np.int64)
    ok = bool(np.all(hk[:-1] <= hk[1:]) and np.array_equal(hk, hk0[hv]))
    same = hk[:-1] == hk[1:]
    ok = ok and bool(np.all(hv[:-1][same] < hv[1:][same]))
    del hk0, hk, hv, same

    if args.calib_copy:
        # A plain copy of known size: 1 GiB read + 1 GiB written per launch.
        src = torch.empty(1 << 28, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ops = {}
    if not args.no_ops:
        del keys0, vals0
        ops = secondary_ops(R, torch, np, dev, stream, n, not args.no_per_segment)

    fa = fm = None
    if not args.no_forall:
        fa = run_example("daxpy_dot_minloc", ["24", "20"])
        if fa:
            for k in ("daxpy", "dot_minloc_reducers", "dot_minloc_params"):
                log("bench: forall %-20s warm %.4f ms (%.1f GB/s)  cold %.4f ms (%.1f GB/s)"
                    % (k, fa[k]["warm_ms"], fa[k]["warm_GBs"], fa[k]["cold_ms"], fa[k]["cold_GBs"]))
        fm = run_example("forall_multi_reducers", [str(args.log2n), "10"])
        if fm:
            for k in ("reducers", "params", "reducers_nontemporal_body"):
                log("bench: forall Sum/Min/Max/MinLoc %-8s %.4f ms (%.1f GB/s, %.1f%% of 8 TB/s)"
                    % (k, fm[k]["ms"], fm[k]["GBs"], 100 * fm[k]["hbm_frac"]))

    halo = None
    if not args.no_halo:
        halo = halo_exchange()

    serve = None
    if not args.no_serve:
        serve = serving(R, torch, np, dev)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(np)

    traffic, traffic_src = committed_traffic(args.log2n)
    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64 counter generator, seed 1; values iota)",
        "config": {
            "workload": "sort_pairs: N=2^%d uint32 keys uniform [0,2^32), int32 iota values, "
                        "ascending, stable LSD onesweep 8-bit digits" % args.log2n,
            "n": n, "passes_executed": p_exec, "bytes_per_elem": bytes_per_elem,
            "hbm_frac": round(gbs / PEAK_GBS, 4), "keys_per_s": round(n / (ms * 1e-3), 1),
            "result_exact": ok, "wall_s_timed_loop": round(wall, 3),
        },
        "roofline": {
            "bound": "hbm", "kernel": "onesweep_pass",
            "achieved": round(pass_gbs, 2), "peak": PEAK_GBS, "unit": "GB/s",
            "frac": round(pass_gbs / PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": n * 16,
            "avg_launch_ms": round(pass_ms, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "hist_kernel_ms": round(hist["total_ms"] / max(1, hist["count"]), 4),
        },
        "cpu_baseline": cpu,
        "shapes": shapes,
        "ops": ops,
        "serving": serve,
        "workgroup_halo": halo,
        "forall_daxpy_dot_minloc": fa,
        "forall_sum_min_max_minloc": fm,
    }
    print(json.dumps(line), flush=True)


def tune_shapes(R, torch, n, measure=True):
    """Shapes of the timed scans and the sort at length n: read from the
    shape cache, measured (raja_cdna4_autotune) where the cache has none."""
    out = {"cache": R.tune_cache_path()}
    for name, op, dt in (("sort_pairs_u32", R.TUNE_SORT_PAIRS, torch.uint32),
                         ("exclusive_scan_i32", R.TUNE_EXCLUSIVE_SCAN, torch.int32),
                         ("inclusive_scan_i64", R.TUNE_INCLUSIVE_SCAN, torch.int64)):
        choice, tuned = R.tune_query(op, dt, n)
        measured = False
        if not tuned and measure:
            try:
                R.autotune(op, dt, n)
                measured = True
            except R.RajaError as e:  # e.g. an unwritable cache directory
                log("bench: shape tuning of %s: %s" % (name, e))
            choice, tuned = R.tune_query(op, dt, n)
        out[name] = {"choice": choice, "tuned": tuned, "measured_now": measured}
        log("bench: shape %s at n=%d: %s (0 = default)" % (name, n, choice))
    return out


TRAFFIC_FILE = os.path.join("profiles", "onesweep_pass_sort_pairs_2p28_hbm_traffic.md")


def committed_traffic(log2n):
    """HBM bytes per onesweep pass at N=2^28, from the committed FETCH_SIZE /
    WRITE_SIZE summary (counter runs scaled by a plain copy of known size)."""
    if log2n != 28:
        return None, None
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            m = re.search(r"^traffic_bytes_per_launch: (\d+)$", f.read(), re.M)
    except OSError:
        return None, None
    return (int(m.group(1)), TRAFFIC_FILE) if m else (None, None)


def secondary_ops(R, torch, np, dev, stream, n, per_segment=True):
    out = {}

    def rec(name, times, nbytes, extra=None):
        ms = statistics.median(times)
        g = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "GB/s": round(g, 1), "hbm_frac": round(g / PEAK_GBS, 4)}
        if extra:
            out[name].update(extra)
        log("bench: %-28s %.4f ms  %.1f GB/s  (%.1f%% of 8 TB/s)" % (name, ms, g, 100 * g / PEAK_GBS))

    x = torch.empty(n, dtype=torch.float64, device=dev)
    R.fill(R.FILL_F64_NORMAL, x, seed=5)
    res = torch.empty(6, dtype=torch.float64, device=dev)
    t = event_time(lambda: R.multi_reduce_f64(x, res), stream, reps=7)
    rec("multi_reduce_f64_sum_min_max_minloc", t, 8 * n)
    t = event_time(lambda: R.reduce(x, R.operators.plus, 0.0, res[:1]), stream, reps=7)
    rec("reduce_sum_f64", t, 8 * n)
    del x, res

    xi = torch.empty(n, dtype=torch.int32, device=dev)
    R.fill(R.FILL_I32, xi, seed=3, a=0, b=16)
    oi = torch.empty_like(xi)
    t = event_time(lambda: R.exclusive_scan(xi, oi, R.operators.plus, 0), stream, reps=7)
    rec("exclusive_scan_i32", t, 8 * n)
    del xi, oi

    xl = torch.empty(n, dtype=torch.int64, device=dev)
    R.fill(R.FILL_I64, xl, seed=4, a=0, b=16)
    ol = torch.empty_like(xl)
    t = event_time(lambda: R.inclusive_scan(xl, ol, R.operators.plus), stream, reps=7)
    rec("inclusive_scan_i64", t, 16 * n)
    del xl, ol

    k0 = torch.empty(n, dtype=torch.uint32, device=dev)
    R.fill(R.FILL_U32_BITS, k0, seed=2, a=12)
    v0 = torch.empty(n, dtype=torch.int32, device=dev)
    R.fill(R.FILL_I32_IOTA, v0, a=0)
    k = torch.empty_like(k0)
    v = torch.empty_like(v0)

    def prep():
        k.copy_(k0)
        v.copy_(v0)
    t = event_time(lambda: R.sort_pairs(k, v), stream, reps=5, prep=prep)
    rec("sort_pairs_u32_dup12", t, n * (16 * 2 + 4),
        {"passes_executed": 2, "bytes_per_elem": 36, "pass_count_bound_full_P4_GBs_equiv":
         round(n * 68 / (statistics.median(t) * 1e-3) / 1e9, 1)})
    del k0, v0, k, v
    out.update(segmented_ops(R, torch, np, dev, stream, rec, per_segment))
    return out


def segmented_ops(R, torch, np, dev, stream, rec, per_segment=True):
    """20000 segments of length clamp(floor(2^(16u)), 1, 65536), u uniform
    (seed 6): batched FP64 sum and int32 exclusive scan over one plan, and
    the same work as one C-ABI call per segment issued from a loop."""
    import ctypes
    nseg = 20000
    u = np.random.default_rng(6).random(nseg)
    lens = np.clip(np.floor(2.0 ** (16 * u)), 1, 1 << 16).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(off[-1])
    plan = R.Segments(off)
    x = torch.empty(total, dtype=torch.float64, device=dev)
    R.fill(R.FILL_F64_NORMAL, x, seed=6)
    sums = torch.empty(nseg, dtype=torch.float64, device=dev)
    xi = torch.empty(total, dtype=torch.int32, device=dev)
    R.fill(R.FILL_I32, xi, seed=6, a=0, b=16)
    oi = torch.empty_like(xi)
    R.segmented_reduce(plan, x, R.operators.plus, 0.0, sums)  # builds the plan's tables
    R.segmented_exclusive_scan(plan, xi, oi, R.operators.plus, 0)
    out = {}
    t = event_time(lambda: R.segmented_reduce(plan, x, R.operators.plus, 0.0, sums), stream, reps=9)
    rec("segmented_sum_f64_20000", t, 8 * total + 8 * nseg, {"elements": total})
    tb = event_time(lambda: R.segmented_exclusive_scan(plan, xi, oi, R.operators.plus, 0), stream, reps=9)
    rec("segmented_exclusive_scan_i32_20000", tb, 8 * total, {"elements": total})

    if not per_segment:
        plan.close()
        return out
    L = R.lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    zd, zi = ctypes.c_double(0.0), ctypes.c_int32(0)
    xp, sp_out, ip, op_ = x.data_ptr(), sums.data_ptr(), xi.data_ptr(), oi.data_ptr()

    def per_seg_sum():
        f = L.raja_cdna4_reduce
        for g in range(nseg):
            f(5, 0, xp + 8 * int(off[g]), int(lens[g]), ctypes.byref(zd), sp_out + 8 * g, sp)

    def per_seg_scan():
        f = L.raja_cdna4_exclusive_scan
        for g in range(nseg):
            f(0, 0, ip + 4 * int(off[g]), op_ + 4 * int(off[g]), int(lens[g]), ctypes.byref(zi), sp)
    t1 = event_time(per_seg_sum, stream, reps=3)
    t2 = event_time(per_seg_scan, stream, reps=3)
    rec("per_segment_calls_sum_f64_20000", t1, 8 * total + 8 * nseg)
    rec("per_segment_calls_exclusive_scan_i32_20000", t2, 8 * total)
    out_speed = {
        "segmented_speedup_vs_per_segment_calls": {
            "sum": round(statistics.median(t1) / statistics.median(t), 2),
            "exclusive_scan": round(statistics.median(t2) / statistics.median(tb), 2)}}
    log("bench: segmented speedup vs per-segment calls: %s" % out_speed)
    plan.close()
    out.update(out_speed)
    return out


def run_example(name, args):
    """Runs raja_amd/bin/<name> (a C++ front-end program) in a child process
    and returns the JSON line it prints."""
    import subprocess
    exe = os.path.join(ROOT, "raja_amd", "bin", name)
    if not os.path.exists(exe):
        log("bench: raja_amd/bin/%s not built; skipped" % name)
        return None
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError("%s failed: %s%s" % (name, r.stdout, r.stderr))
    return json.loads(r.stdout.strip().splitlines()[-1])


def halo_exchange():
    """The WorkGroup halo exchange (256^3 grid, halo 2, 10 FP64 variables,
    260 loops per phase) run by raja_amd/bin/halo_exchange in a child
    process: per-loop forall, ordered WorkGroup, unordered tile WorkGroup."""
    out = run_example("halo_exchange", ["256", "256", "256", "2", "10", "5"])
    if out is None:
        return None
    ph = lambda v, k: out[v]["pack_%s_ms" % k] + out[v]["unpack_%s_ms" % k]
    out["speedup_unordered_vs_ordered_device"] = round(ph("ordered", "device") / ph("unordered", "device"), 3)
    out["speedup_unordered_vs_forall_host"] = round(ph("forall", "host") / ph("unordered", "host"), 3)
    for v in ("forall", "ordered", "unordered"):
        log("bench: halo %-9s pack %.4f ms unpack %.4f ms (device), %.4f + %.4f ms (host), %.1f / %.1f GB/s"
            % (v, out[v]["pack_device_ms"], out[v]["unpack_device_ms"], out[v]["pack_host_ms"],
               out[v]["unpack_host_ms"], out[v]["pack_GBs"], out[v]["unpack_GBs"]))
    return out


def serve_mix(np, nreq, seed, lo_log2, span_log2):
    """Request kinds 40% FP64 sum / 30% int32 exclusive scan / 30% sort_pairs,
    sizes floor(2^(lo + span*u))."""
    rng = np.random.default_rng(seed)
    u = rng.random(nreq)
    kinds = np.where(u < 0.4, 0, np.where(u < 0.7, 1, 2))
    sizes = np.floor(2.0 ** (lo_log2 + span_log2 * rng.random(nreq))).astype(np.int64)
    return kinds, sizes, rng


def serving(R, torch, np, dev, nreq=4000):
    """The request stream: saturation req/s (every request waiting at t=0),
    batched vs one launch sequence per request, and p50/p99 latency under
    Poisson arrivals below the batched saturation rate."""
    nmax, ring = 1 << 24, 1 << 28
    x = torch.empty(nmax, dtype=torch.float64, device=dev)
    R.fill(R.FILL_F64_NORMAL, x, seed=5)
    si = torch.empty(nmax, dtype=torch.int32, device=dev)
    R.fill(R.FILL_I32, si, seed=3, a=0, b=16)
    sk = torch.empty(nmax, dtype=torch.uint32, device=dev)
    R.fill(R.FILL_U32_BITS, sk, seed=1, a=32)
    sv = torch.arange(nmax, dtype=torch.int32, device=dev)
    sums = torch.empty(nreq, dtype=torch.float64, device=dev)
    so = torch.empty(ring, dtype=torch.int32, device=dev)
    ok_ = torch.empty(ring, dtype=torch.uint32, device=dev)
    ov = torch.empty(ring, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def build(kinds, sizes, rng, arrivals=None):
        reqs = (R.Request * len(kinds))()
        cur = [0, 0]
        nbytes = 0
        for i, (k, n) in enumerate(zip(kinds, sizes)):
            n = int(n)
            src = int(rng.integers(0, (nmax - n) // 64 + 1)) * 64
            t = 0.0 if arrivals is None else float(arrivals[i])
            if k == 0:
                reqs[i] = R.request(0, n, x.data_ptr() + 8 * src, sums.data_ptr() + 8 * i, arrival_s=t)
                nbytes += 8 * n
                continue
            c = 0 if k == 1 else 1
            if cur[c] + n > ring:
                cur[c] = 0
            o = cur[c]
            cur[c] += (n + 63) // 64 * 64
            if k == 1:
                reqs[i] = R.request(1, n, si.data_ptr() + 4 * src, so.data_ptr() + 4 * o, arrival_s=t)
                nbytes += 8 * n
            else:
                reqs[i] = R.request(2, n, sk.data_ptr() + 4 * src, ok_.data_ptr() + 4 * o,
                                    sv.data_ptr() + 4 * src, ov.data_ptr() + 4 * o, arrival_s=t)
                nbytes += 68 * n
        return reqs, nbytes

    def lat(res, arrivals):
        d = np.array([r.done_s for r in res]) - np.asarray(arrivals)
        return float(np.percentile(d, 50) * 1e3), float(np.percentile(d, 99) * 1e3)

    # Every request is batchable (batch_limit 2^24, the largest request):
    # units of up to 512 waiting requests, one kernel (or one per digit pass)
    # per kind over all of them.
    cfg = dict(streams=16, max_inflight=64, max_batch=512, batch_limit=1 << 24, arena_bytes=64 << 30)
    out = {"workload": "%d requests: 40%% FP64 sum, 30%% int32 exclusive scan, 30%% sort_pairs "
                       "(uint32, int32); n = floor(2^(10+14u)); inputs resident; 64 GB arena" % nreq}
    servers = {
        "batched": R.Server(R.SERVE_BATCHED, **cfg),
        "per_request": R.Server(R.SERVE_PER_REQUEST, 1, 8, 1, 0, 64 << 30),
    }
    kinds, sizes, rng = serve_mix(np, nreq, 7, 10, 14)
    reqs, nbytes = build(kinds, sizes, rng)
    sat = {}
    for name, srv in servers.items():
        srv.run(reqs[:200])  # warm-up
        walls = []
        for _ in range(3):
            res, wall = srv.run(reqs)
            assert all(r.status == 0 for r in res)
            walls.append(wall)
        w = statistics.median(walls)
        sat[name] = nreq / w
        out["saturation_" + name] = {"req_per_s": round(nreq / w, 1), "wall_s": round(w, 4),
                                     "GB_per_s": round(nbytes / w / 1e9, 1)}
        log("bench: serve saturation %-12s %9.1f req/s  %.1f GB/s" % (name, nreq / w, nbytes / w / 1e9))
    out["saturation_speedup"] = round(sat["batched"] / sat["per_request"], 3)

    # small-request stream (n in [2^10, 2^16]): launch-latency bound
    kinds_s, sizes_s, rng_s = serve_mix(np, nreq, 8, 10, 6)
    reqs_s, nbytes_s = build(kinds_s, sizes_s, rng_s)
    sat_s = {}
    for name, srv in servers.items():
        srv.run(reqs_s[:200])
        walls = [srv.run(reqs_s)[1] for _ in range(3)]
        w = statistics.median(walls)
        sat_s[name] = nreq / w
        out["small_saturation_" + name] = {"req_per_s": round(nreq / w, 1), "wall_s": round(w, 4)}
        log("bench: serve small-request saturation %-12s %9.1f req/s" % (name, nreq / w))
    out["small_saturation_speedup"] = round(sat_s["batched"] / sat_s["per_request"], 3)

    # Poisson arrivals at 60% of the batched saturation rate
    lam = 0.6 * sat["batched"]
    arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / lam, nreq))
    reqs_p, _ = build(kinds, sizes, np.random.default_rng(9), arr)
    for name, args_ in (("batched", (R.SERVE_BATCHED, 4, 8, 512, 1 << 24, 64 << 30, True)),
                        ("per_request", (R.SERVE_PER_REQUEST, 1, 8, 1, 0, 64 << 30, True))):
        servers[name].close()
        srv = R.Server(*args_)
        res, wall = srv.run(reqs_p)
        p50, p99 = lat(res, arr)
        out["poisson_" + name] = {"offered_req_per_s": round(lam, 1), "req_per_s": round(nreq / wall, 1),
                                  "p50_ms": round(p50, 4), "p99_ms": round(p99, 4)}
        log("bench: serve poisson %.0f req/s %-12s p50 %.3f ms  p99 %.3f ms" % (lam, name, p50, p99))
        srv.close()

    # The batched server's latency over a sweep of offered rates (fractions
    # of its saturation rate): where p99 turns up is the knee.
    srv = R.Server(R.SERVE_BATCHED, 4, 8, 512, 1 << 24, 64 << 30, True)
    sweep = []
    for frac in (0.2, 0.4, 0.6, 0.8, 0.9, 1.0):
        lam_f = frac * sat["batched"]
        arr_f = np.cumsum(np.random.default_rng(11).exponential(1.0 / lam_f, nreq))
        reqs_f, _ = build(kinds, sizes, np.random.default_rng(9), arr_f)
        res, wall = srv.run(reqs_f)
        assert all(r.status == 0 for r in res)
        p50, p99 = lat(res, arr_f)
        sweep.append({"fraction_of_saturation": frac, "offered_req_per_s": round(lam_f, 1),
                      "req_per_s": round(nreq / wall, 1), "p50_ms": round(p50, 4), "p99_ms": round(p99, 4)})
        log("bench: serve poisson sweep %.1f x saturation (%.0f req/s) batched p50 %.3f ms  p99 %.3f ms"
            % (frac, lam_f, p50, p99))
    srv.close()
    out["poisson_batched_sweep"] = sweep
    return out


def _host_info(O):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "omp_max_threads": O.omp_max_threads(),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def _cpu_time(fn, reps=5, prep=None):
    times = []
    for _ in range(reps):
        if prep is not None:
            prep()
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline(np):
    """The oracle's omp_parallel_for_exec restatements (RAJA's OpenMP
    backend semantics) timed on this host's cores, median of 5, on bounded
    samples of each workload; GB/s uses the GPU's algorithmic byte counts.
    The headline entry is the stable sort_pairs."""
    from oracle import oracle as O
    threads = O.omp_max_threads()
    host = _host_info(O)
    ops = {}

    def rec(name, sec, nbytes, sample):
        ops[name] = {"s": round(sec, 5), "GB/s": round(nbytes / sec / 1e9, 3), "sample": sample}
        log("bench: cpu omp %-26s %.4f s  %.2f GB/s  (%s)" % (name, sec, nbytes / sec / 1e9, sample))

    m = 1 << 26
    keys = O.fill_u32_bits(m, 1, 32)
    vals = np.arange(m, dtype=np.int32)
    k, v = keys.copy(), vals.copy()

    def prep():
        k[:] = keys
        v[:] = vals
    s = _cpu_time(lambda: O.omp_stable_sort_pairs(k, v, nthreads=threads), prep=prep)
    gbs = m * 68 / s / 1e9
    log("bench: cpu omp stable_sort_pairs 2^26: %.3f s (%.2f GB/s equiv, %d threads)" % (s, gbs, threads))
    del keys, vals, k, v

    x = O.fill_f64_normal(m, 5)
    rec("multi_reduce_f64_sum_min_max_minloc", _cpu_time(lambda: O.omp_multi_reduce_f64(x, threads)),
        8 * m, "2^26 N(0,1) FP64, seed 5")
    rec("reduce_sum_f64", _cpu_time(lambda: O.omp_sum_f64(x, threads)), 8 * m, "2^26 N(0,1) FP64")
    del x
    xi = O.fill_i32_range(m, 3, 0, 16)
    oi = np.empty_like(xi)
    rec("exclusive_scan_i32", _cpu_time(lambda: O.omp_exclusive_i32(xi, oi, 0, threads)), 8 * m,
        "2^26 int32 in [0,16), seed 3")
    del xi, oi
    xl = O.fill_i64_range(m, 4, 0, 16)
    ol = np.empty_like(xl)
    rec("inclusive_scan_i64", _cpu_time(lambda: O.omp_inclusive_i64(xl, ol, threads)), 16 * m,
        "2^26 int64 in [0,16), seed 4")
    del xl, ol

    nd = 1 << 24
    xd = O.fill_f64_uniform(nd, 0, -1.0, 1.0)
    yd = O.fill_f64_uniform(nd, 10, -1.0, 1.0)
    rec("daxpy_dot_minloc_f64", _cpu_time(lambda: O.omp_daxpy_dot_minloc(xd, yd, 2.5, threads)),
        (24 + 16) * nd, "2^24 FP64 (the full forall example size): daxpy then dot + MinLoc")
    del xd, yd

    g, h, nv = 256, 2, 10
    plen, plists = O.halo_lists([g, g, g], h, False)
    _, ulists = O.halo_lists([g, g, g], h, True)
    cells = (g + 2 * h) ** 3
    vars_ = [np.arange(cells, dtype=np.float64) + vi for vi in range(nv)]
    buf = np.empty(int(plen.sum()) * nv, np.float64)
    rec("halo_pack_unpack_cycle", _cpu_time(lambda: O.omp_halo_cycle(vars_, plists, ulists, plen, buf,
                                                                      threads)),
        2 * 20 * int(plen.sum()) * nv, "256^3 grid, halo 2, 10 FP64 variables (the full size)")
    del vars_, buf

    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": "stable sort_pairs of 2^26 uint32 keys uniform [0,2^32) + int32 iota "
                      "(same generator, seed 1); median of 5; GB/s = 68 B/elem x N / time, "
                      "the GPU's algorithmic byte count",
            "seconds": round(s, 3), "keys_per_s": round(m / s, 1), "host": host, "ops": ops}


if __name__ == "__main__":
    main()
