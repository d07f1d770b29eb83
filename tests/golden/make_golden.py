"""Write tests/golden/golden.json: known-answer vectors for the hot path.

The reference stores no golden data for this path; its tests compute their
expected values inline.  This script restates those inline cases with
closed forms and library routines (pure Python / numpy, no use of the
oracle), so the fixtures can pin both the oracle and the GPU path:

* rng: the published SplitMix64 first output for state 0, and the counter
  stream (seed << 40) + i for a few seeds, from a pure-Python SplitMix64;
* scan: iota(1..N) inputs for N in {0, 357, 32000} with offsets
  {identity, 13, 15, 2} and ops {plus, minimum, maximum}
  (test/functional/scan/tests/test-scan-*.hpp; test-scan.cpp.in) -- expected
  values from the recurrences' closed forms (triangular numbers, running
  min/max of an increasing sequence);
* sort: keys uniform in [-N, N] (duplicates, negatives) for the sizes of
  test/unit/algorithm/tests/test-algorithm-sort-utils.hpp:610-706, with
  iota values; expected stable order from Python's sorted() (stable), for
  less and greater;
* minloc/maxloc: planted unique extremum at len*2/3, all-equal arrays
  (first index wins), init that dominates (loc -1)
  (test/functional/forall/reduce-basic/tests/test-forall-basic-ReduceMinLoc.hpp);
* reduce sum: rand()%100-style integer data with init offsets
  (test-forall-basic-ReduceSum.hpp) -- exact sums;
* halo: list lengths and first/last index of each of the 26 pack and unpack
  lists for a 256^3 grid with halo 2 (examples/tut_halo-exchange.cpp:1793-2062),
  from the extents' closed form.
"""
import json
import os
import random

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
M64 = (1 << 64) - 1


def splitmix64_output(state):
    z = (state + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def rng_cases():
    cases = []
    for seed in (0, 1, 5, 7):
        raw = [splitmix64_output(((seed << 40) + i) & M64) for i in range(16)]
        cases.append({"seed": seed, "raw": [str(r) for r in raw],
                      "u32_bits32": [r >> 32 for r in raw],
                      "u32_bits12": [r >> 52 for r in raw]})
    return {"splitmix64_state0_first": str(0xE220A8397B1DCDAF), "streams": cases}


def scan_cases():
    out = []
    big = np.iinfo(np.int64).max
    small = np.iinfo(np.int64).min
    for n in (0, 357, 32000):
        x = np.arange(1, n + 1, dtype=np.int64)
        i = np.arange(n, dtype=np.int64)
        for op in ("plus", "minimum", "maximum"):
            ident = {"plus": 0, "minimum": big, "maximum": small}[op]
            # inclusive: closed forms of the recurrence on 1..N
            if op == "plus":
                inc = (i + 1) * (i + 2) // 2
            elif op == "minimum":
                inc = np.ones(n, np.int64)
            else:
                inc = i + 1
            out.append({"n": n, "kind": "inclusive", "op": op, "init": None,
                        "expect": _compact(inc)})
            for init in (None, 13, 15, 2):
                v = ident if init is None else init
                if op == "plus":
                    exc = v + i * (i + 1) // 2
                elif op == "minimum":
                    exc = np.where(i == 0, v, np.minimum(v, 1))
                else:
                    exc = np.where(i == 0, v, np.maximum(v, i))
                out.append({"n": n, "kind": "exclusive", "op": op, "init": init,
                            "expect": _compact(exc)})
    return out


def _compact(a):
    """Full array for short cases, every 97th element plus the last otherwise."""
    a = np.asarray(a, dtype=np.int64)
    if a.size <= 400:
        return {"idx": list(range(a.size)), "val": [int(v) for v in a]}
    idx = list(range(0, a.size, 97)) + [a.size - 1]
    return {"idx": idx, "val": [int(a[j]) for j in idx]}


def sort_cases():
    rnd = random.Random(20250901)
    out = []
    for n in (0, 1, 10, 100, 1000, 10000, rnd.randint(5001, 10000)):
        keys = [rnd.randint(-n, n) for _ in range(n)]
        case = {"n": n, "keys": keys}
        # stable ascending, and stable descending (larger keys first, ties in
        # input order); values are iota, so the expected values are the order.
        case["less"] = sorted(range(n), key=lambda j: keys[j])
        case["greater"] = sorted(range(n), key=lambda j: -keys[j])
        out.append(case)
    # +-0.0 and duplicates in double keys: -0.0 == +0.0 under operator<, so a
    # stable sort keeps them in input order (Python's sorted compares the same)
    fk = [0.0, -0.0, 1.5, -2.25, 0.0, -0.0, 1.5, float("-inf"), float("inf"), -2.25]
    order = sorted(range(len(fk)), key=lambda j: fk[j])
    out.append({"n": len(fk), "keys": fk, "float": True, "less": order})
    return out


def loc_cases():
    rnd = random.Random(7)
    out = []
    for n in (28, 642, 2057, 10567):
        data = [rnd.randint(0, 99) + 10 for _ in range(n)]
        p = n * 2 // 3
        data[p] = -5
        q = n // 3
        data2 = list(data)
        data2[q] = 500
        out.append({"n": n, "data": data, "min_val": -5, "min_loc": p,
                    "data_max": data2, "max_val": 500, "max_loc": q})
    out.append({"n": 1000, "data": [3] * 1000, "min_val": 3, "min_loc": 0,
                "data_max": [3] * 1000, "max_val": 3, "max_loc": 0})
    return out


def sum_cases():
    rnd = random.Random(11)
    out = []
    for n in (28, 639, 2057, 65536 + 3):
        data = [rnd.randint(0, 99) for _ in range(n)]
        out.append({"n": n, "data": data, "inits": [0, 2],
                    "expect": [init + sum(data) for init in (0, 2)]})
    return out


def halo_cases():
    g, h = 256, 2
    dirs = ([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)] +
            [(-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0), (-1, 0, -1), (-1, 0, 1),
             (1, 0, -1), (1, 0, 1), (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1)] +
            [(-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1), (1, -1, 1),
             (1, 1, -1), (1, 1, 1)])
    sj = g + 2 * h
    sk = sj * (g + 2 * h)

    def rng(d, unpack):
        if d == 0:
            return h, g + h
        if d < 0:
            return (0, h) if unpack else (h, 2 * h)
        return (g + h, g + 2 * h) if unpack else (g, g + h)

    res = {"grid": g, "halo": h}
    for unpack in (0, 1):
        lens, first, last = [], [], []
        for d in dirs:
            r = [rng(c, unpack) for c in d]
            lens.append((r[0][1] - r[0][0]) * (r[1][1] - r[1][0]) * (r[2][1] - r[2][0]))
            first.append(r[0][0] + r[1][0] * sj + r[2][0] * sk)
            last.append((r[0][1] - 1) + (r[1][1] - 1) * sj + (r[2][1] - 1) * sk)
        res["unpack" if unpack else "pack"] = {"lengths": lens, "first": first, "last": last}
    return res


def main():
    data = {"rng": rng_cases(), "scan": scan_cases(), "sort": sort_cases(),
            "loc": loc_cases(), "sum": sum_cases(), "halo": halo_cases()}
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
