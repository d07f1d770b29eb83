"""TEST INFRASTRUCTURE ONLY: numpy bindings of oracle/liboracle.so, the CPU
restatement of RAJA's seq_exec / omp_parallel_for_exec semantics.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module.  Each wrapper names the reference file:line whose behaviour the C
function restates (see oracle/raja_oracle.c for the pinning of each group).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_SFX = {np.dtype(np.int32): "i32", np.dtype(np.int64): "i64",
        np.dtype(np.uint32): "u32", np.dtype(np.uint64): "u64",
        np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}
_CT = {"i32": ctypes.c_int32, "i64": ctypes.c_int64, "u32": ctypes.c_uint32,
       "u64": ctypes.c_uint64, "f32": ctypes.c_float, "f64": ctypes.c_double}

PLUS, MIN, MAX = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_rng_raw.restype = ctypes.c_uint64
        _lib.orc_rng_raw.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _lib.orc_omp_max_threads.restype = ctypes.c_int
        _lib.orc_halo_lists.restype = ctypes.c_int64
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fn(name, restype=None):
    f = getattr(lib(), name)
    f.restype = restype
    return f


# ---- generator ------------------------------------------------------------------

def rng_raw(seed, i):
    return lib().orc_rng_raw(seed, i)


def fill_u32_bits(n, seed, bits):
    out = np.empty(n, np.uint32)
    _fn("orc_fill_u32_bits")(_p(out), ctypes.c_int64(n), ctypes.c_uint64(seed), ctypes.c_int(bits))
    return out


def fill_i32_range(n, seed, lo, hi):
    out = np.empty(n, np.int32)
    _fn("orc_fill_i32_range")(_p(out), ctypes.c_int64(n), ctypes.c_uint64(seed),
                              ctypes.c_int32(lo), ctypes.c_int32(hi))
    return out


def fill_i64_range(n, seed, lo, hi):
    out = np.empty(n, np.int64)
    _fn("orc_fill_i64_range")(_p(out), ctypes.c_int64(n), ctypes.c_uint64(seed),
                              ctypes.c_int64(lo), ctypes.c_int64(hi))
    return out


def fill_f64_uniform(n, seed, a, b):
    out = np.empty(n, np.float64)
    _fn("orc_fill_f64_uniform")(_p(out), ctypes.c_int64(n), ctypes.c_uint64(seed),
                                ctypes.c_double(a), ctypes.c_double(b))
    return out


def fill_f64_normal(n, seed):
    out = np.empty(n, np.float64)
    _fn("orc_fill_f64_normal")(_p(out), ctypes.c_int64(n), ctypes.c_uint64(seed))
    return out


# ---- reducers (include/RAJA/pattern/detail/reduce.hpp:282-593) --------------------

def seq_sum(x, init=0):
    x = np.ascontiguousarray(x)
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    return _fn("orc_seq_sum_" + sfx, ct)(_p(x), ctypes.c_int64(x.size), ct(init))


def seq_dot_f64(x, y, init=0.0):
    return _fn("orc_seq_dot_f64", ctypes.c_double)(_p(x), _p(y), ctypes.c_int64(x.size),
                                                  ctypes.c_double(init))


def kn_sum_f64(x):
    s, a = ctypes.c_double(), ctypes.c_double()
    _fn("orc_kn_sum_f64")(_p(x), ctypes.c_int64(x.size), ctypes.byref(s), ctypes.byref(a))
    return s.value, a.value


def kn_dot_f64(x, y):
    s, a = ctypes.c_double(), ctypes.c_double()
    _fn("orc_kn_dot_f64")(_p(x), _p(y), ctypes.c_int64(x.size), ctypes.byref(s), ctypes.byref(a))
    return s.value, a.value


def _ident_min(dt):
    return np.finfo(dt).max if dt.kind == "f" else np.iinfo(dt).max


def _ident_max(dt):
    return -np.finfo(dt).max if dt.kind == "f" else np.iinfo(dt).min


def seq_min(x, init=None):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    ident = _ident_min(x.dtype)
    init = ident if init is None else init
    return _fn("orc_seq_min_" + sfx, ct)(_p(x), ctypes.c_int64(x.size), ct(init), ct(ident))


def seq_max(x, init=None):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    ident = _ident_max(x.dtype)
    init = ident if init is None else init
    return _fn("orc_seq_max_" + sfx, ct)(_p(x), ctypes.c_int64(x.size), ct(init), ct(ident))


def seq_minloc(x, init=None, init_loc=-1):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    ident = _ident_min(x.dtype)
    init = ident if init is None else init
    v, l = ct(), ctypes.c_int64()
    _fn("orc_seq_minloc_" + sfx)(_p(x), ctypes.c_int64(x.size), ct(init), ctypes.c_int64(init_loc),
                                 ct(ident), ctypes.byref(v), ctypes.byref(l))
    return v.value, l.value


def seq_maxloc(x, init=None, init_loc=-1):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    ident = _ident_max(x.dtype)
    init = ident if init is None else init
    v, l = ct(), ctypes.c_int64()
    _fn("orc_seq_maxloc_" + sfx)(_p(x), ctypes.c_int64(x.size), ct(init), ctypes.c_int64(init_loc),
                                 ct(ident), ctypes.byref(v), ctypes.byref(l))
    return v.value, l.value


def seq_bitor(x, init=0):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    return _fn("orc_seq_bitor_" + sfx, ct)(_p(x), ctypes.c_int64(x.size), ct(init))


def seq_bitand(x, init=-1):
    sfx = _SFX[x.dtype]
    ct = _CT[sfx]
    return _fn("orc_seq_bitand_" + sfx, ct)(_p(x), ctypes.c_int64(x.size), ct(init))


# ---- scans (include/RAJA/policy/sequential/scan.hpp:43-153) ------------------------

def seq_inclusive(x, op=PLUS):
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    _fn("orc_seq_inclusive_" + _SFX[x.dtype])(_p(x), _p(out), ctypes.c_int64(x.size), ctypes.c_int(op))
    return out


def seq_exclusive(x, op=PLUS, init=None):
    x = np.ascontiguousarray(x)
    sfx = _SFX[x.dtype]
    if init is None:
        init = {PLUS: 0, MIN: _ident_min(x.dtype), MAX: _ident_max(x.dtype)}[op]
    out = np.empty_like(x)
    _fn("orc_seq_exclusive_" + sfx)(_p(x), _p(out), ctypes.c_int64(x.size), ctypes.c_int(op),
                                    _CT[sfx](init))
    return out


# ---- sorts (include/RAJA/util/sort.hpp:576-693 merge_sort, stable) --------------

def seq_stable_sort_pairs(keys, vals=None, descending=False):
    k = np.array(keys, copy=True)
    v = None if vals is None else np.array(vals, copy=True)
    vb = 0 if v is None else v.itemsize
    _fn("orc_seq_stable_sort_pairs_" + _SFX[k.dtype])(
        _p(k), None if v is None else _p(v), ctypes.c_int64(k.size),
        ctypes.c_int(1 if descending else 0), ctypes.c_int(vb))
    return k, v


def omp_stable_sort_pairs(keys, vals, descending=False, nthreads=None):
    nthreads = nthreads or lib().orc_omp_max_threads()
    _fn("orc_omp_stable_sort_pairs_" + _SFX[keys.dtype])(
        _p(keys), _p(vals), ctypes.c_int64(keys.size), ctypes.c_int(1 if descending else 0),
        ctypes.c_int(vals.itemsize), ctypes.c_int(nthreads))
    return keys, vals


# ---- omp baselines (include/RAJA/policy/openmp/*) ------------------------------------

def omp_max_threads():
    return lib().orc_omp_max_threads()


def omp_multi_reduce_f64(x, nthreads=None):
    nthreads = nthreads or omp_max_threads()
    r = [ctypes.c_double() for _ in range(4)]
    loc = ctypes.c_int64()
    _fn("orc_omp_multi_reduce_f64")(_p(x), ctypes.c_int64(x.size), ctypes.c_int(nthreads),
                                    *[ctypes.byref(v) for v in r], ctypes.byref(loc))
    return [v.value for v in r] + [loc.value]


def omp_sum_f64(x, nthreads=None):
    """ReduceSum<omp_reduce, double> over one forall
    (include/RAJA/policy/openmp/reduce.hpp:46-69)."""
    nthreads = nthreads or omp_max_threads()
    return _fn("orc_omp_sum_f64", ctypes.c_double)(_p(x), ctypes.c_int64(x.size),
                                                  ctypes.c_int(nthreads))


def omp_exclusive_i32(x, out, init=0, nthreads=None):
    nthreads = nthreads or omp_max_threads()
    _fn("orc_omp_exclusive_i32")(_p(x), _p(out), ctypes.c_int64(x.size), ctypes.c_int32(init),
                                 ctypes.c_int(nthreads))
    return out


def omp_inclusive_i64(x, out, nthreads=None):
    nthreads = nthreads or omp_max_threads()
    _fn("orc_omp_inclusive_i64")(_p(x), _p(out), ctypes.c_int64(x.size), ctypes.c_int(nthreads))
    return out


def omp_daxpy_dot_minloc(x, y, a, nthreads=None):
    nthreads = nthreads or omp_max_threads()
    d, lv, ll = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    _fn("orc_omp_daxpy_dot_minloc")(_p(x), _p(y), ctypes.c_int64(x.size), ctypes.c_double(a),
                                    ctypes.c_int(nthreads), ctypes.byref(d), ctypes.byref(lv),
                                    ctypes.byref(ll))
    return d.value, lv.value, ll.value


# ---- segmented ------------------------------------------------------------------

def seg_sum_f64(x, offsets):
    out = np.empty(offsets.size - 1, np.float64)
    _fn("orc_seg_sum_f64")(_p(x), _p(offsets), ctypes.c_int64(offsets.size - 1), _p(out))
    return out


def seg_kn_sum_f64(x, offsets):
    out = np.empty(offsets.size - 1, np.float64)
    ab = np.empty(offsets.size - 1, np.float64)
    _fn("orc_seg_kn_sum_f64")(_p(x), _p(offsets), ctypes.c_int64(offsets.size - 1), _p(out), _p(ab))
    return out, ab


def seg_exclusive_i32(x, offsets, init=0):
    out = np.empty_like(x)
    _fn("orc_seg_exclusive_i32")(_p(x), _p(out), _p(offsets), ctypes.c_int64(offsets.size - 1),
                                 ctypes.c_int32(init))
    return out


def seg_stable_sort_pairs_u32(keys, vals, offsets, descending=False):
    k = keys.copy()
    v = vals.copy()
    _fn("orc_seg_stable_sort_pairs_u32")(_p(k), _p(v), _p(offsets), ctypes.c_int64(offsets.size - 1),
                                         ctypes.c_int(1 if descending else 0), ctypes.c_int(v.itemsize))
    return k, v


# ---- halo exchange (examples/tut_halo-exchange.cpp:1793-2062) ---------------------

def halo_lists(grid_dims, halo, unpack):
    g = np.asarray(grid_dims, np.int32)
    lengths = np.empty(26, np.int32)
    total = lib().orc_halo_lists(_p(g), ctypes.c_int(halo), ctypes.c_int(1 if unpack else 0),
                                 _p(lengths), None)
    lists = np.empty(total, np.int32)
    lib().orc_halo_lists(_p(g), ctypes.c_int(halo), ctypes.c_int(1 if unpack else 0), _p(lengths),
                         _p(lists))
    return lengths, lists


def halo_cycle(vars_, pack_lists, unpack_lists, lengths):
    nv = len(vars_)
    ptrs = (ctypes.c_void_p * nv)(*[v.ctypes.data for v in vars_])
    buf = np.empty(int(lengths.sum()) * nv, np.float64)
    _fn("orc_halo_cycle")(ptrs, ctypes.c_int(nv), _p(pack_lists), _p(unpack_lists), _p(lengths),
                          _p(buf))
    return vars_


def omp_halo_cycle(vars_, pack_lists, unpack_lists, lengths, buf=None, nthreads=None):
    """The example's omp_parallel_for_exec halo cycle
    (examples/tut_halo-exchange.cpp:615-700): one parallel loop per
    (neighbour, variable), packs then unpacks."""
    nthreads = nthreads or omp_max_threads()
    nv = len(vars_)
    ptrs = (ctypes.c_void_p * nv)(*[v.ctypes.data for v in vars_])
    if buf is None:
        buf = np.empty(int(lengths.sum()) * nv, np.float64)
    _fn("orc_omp_halo_cycle")(ptrs, ctypes.c_int(nv), _p(pack_lists), _p(unpack_lists),
                              _p(lengths), _p(buf), ctypes.c_int(nthreads))
    return vars_
