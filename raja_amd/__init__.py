"""raja_amd -- Python bindings of libraja_cdna4.so, the MI355X execution library
behind the RAJA::cdna4_* policy family.

The names mirror RAJA's pattern interface for the hot path
(include/RAJA/pattern/scan.hpp, include/RAJA/pattern/sort.hpp,
include/RAJA/pattern/reduce.hpp): ``inclusive_scan(in, out, op)``,
``exclusive_scan(in, out, op, init)``, the ``*_inplace`` variants,
``sort/stable_sort/sort_pairs/stable_sort_pairs(keys[, vals], comparator)``
and span reductions.  Arguments are torch tensors on a GPU (torch is used for
device memory and streams only); every call goes through the C ABI declared
in include/raja_cdna4.h and runs a hand-written gfx950 kernel.  There is no
fallback: if the library is missing or the tensor is not on a GPU, the call
raises.
"""
import ctypes
import os

import torch  # loads the HIP runtime the library then binds to (same soname)

__all__ = [
    "operators", "lib", "RajaError", "inclusive_scan", "exclusive_scan",
    "inclusive_scan_inplace", "exclusive_scan_inplace", "sort", "stable_sort",
    "sort_pairs", "stable_sort_pairs", "reduce", "reduce_sum", "reduce_min",
    "reduce_max", "reduce_minloc", "reduce_maxloc", "multi_reduce_f64", "fill",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libraja_cdna4.so")

# enum raja_cdna4_dtype / raja_cdna4_op
_DTYPES = {
    torch.int32: 0, torch.int64: 1, torch.uint32: 2, torch.uint64: 3,
    torch.float32: 4, torch.float64: 5,
}


class operators:
    """RAJA::operators tags (include/RAJA/util/Operators.hpp)."""
    plus = 0
    minimum = 1
    maximum = 2
    bit_or = 3
    bit_and = 4
    less = "less"
    greater = "greater"


class RajaError(RuntimeError):
    pass


_lib = None

# ---- request serving (raja_cdna4_serve_config / _request / _result) --------------

REQ_SUM_F64 = 0
REQ_EXCLUSIVE_SCAN_I32 = 1
REQ_SORT_PAIRS_U32_I32 = 2
SERVE_BATCHED = 0
SERVE_PER_REQUEST = 1
STATUS_ADMISSION = 5


class ServeConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("streams", ctypes.c_int),
                ("max_inflight", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("batch_limit", ctypes.c_int64), ("arena_bytes", ctypes.c_int64),
                ("honor_arrivals", ctypes.c_int), ("pad", ctypes.c_int),
                ("batch_bytes", ctypes.c_int64)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("pad", ctypes.c_int), ("n", ctypes.c_int64),
                ("in_", ctypes.c_void_p), ("in_vals", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("out_vals", ctypes.c_void_p),
                ("arrival_s", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("unit", ctypes.c_int), ("unit_size", ctypes.c_int),
                ("pad", ctypes.c_int), ("start_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_int64), ("result", Result)]


def _declare(L):
    vp, i64, i32, u64, dbl = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                              ctypes.c_uint64, ctypes.c_double)
    sigs = {
        "raja_cdna4_init": [i32],
        "raja_cdna4_num_cus": [],
        "raja_cdna4_stream_synchronize": [vp],
        "raja_cdna4_stream_release": [vp],
        "raja_cdna4_device_status": [],
        "raja_cdna4_fill": [i32, vp, i64, u64, dbl, dbl, vp],
        "raja_cdna4_reduce": [i32, i32, vp, i64, vp, vp, vp],
        "raja_cdna4_reduce_loc": [i32, i32, vp, i64, vp, i64, vp, vp, vp],
        "raja_cdna4_multi_reduce_f64": [vp, i64, vp, vp],
        "raja_cdna4_inclusive_scan": [i32, i32, vp, vp, i64, vp],
        "raja_cdna4_exclusive_scan": [i32, i32, vp, vp, i64, vp, vp],
        "raja_cdna4_exclusive_scan_noalloc": [i32, i32, vp, vp, i64, vp, vp],
        "raja_cdna4_sort_pairs": [i32, vp, vp, i32, i64, i32, vp],
        "raja_cdna4_autotune": [i32, i32, i64, vp],
        "raja_cdna4_tune_query": [i32, i32, i64, ctypes.POINTER(i32), ctypes.POINTER(i32)],
        "raja_cdna4_tune_cache_path": [ctypes.c_char_p, ctypes.c_size_t],
        "raja_cdna4_tune_forget": [],
        "raja_cdna4_timing_enable": [i32],
        "raja_cdna4_timing_reset": [],
        "raja_cdna4_timing_query": [ctypes.c_char_p, ctypes.POINTER(i64),
                                    ctypes.POINTER(dbl), ctypes.POINTER(dbl),
                                    ctypes.POINTER(dbl)],
        "raja_cdna4_timing_names": [ctypes.c_char_p, ctypes.c_size_t],
        "raja_cdna4_segments_create": [vp, i64, ctypes.POINTER(vp)],
        "raja_cdna4_segments_destroy": [vp],
        "raja_cdna4_segmented_reduce": [vp, i32, i32, vp, vp, vp, vp],
        "raja_cdna4_segmented_exclusive_scan": [vp, i32, i32, vp, vp, vp, vp],
        "raja_cdna4_segmented_inclusive_scan": [vp, i32, i32, vp, vp, vp],
        "raja_cdna4_segmented_sort_pairs": [vp, i32, vp, vp, i32, i32, vp],
        "raja_cdna4_server_create": [ctypes.POINTER(ServeConfig), ctypes.POINTER(vp)],
        "raja_cdna4_server_run": [vp, ctypes.POINTER(Request), i64, ctypes.POINTER(Result),
                                  ctypes.POINTER(dbl)],
        "raja_cdna4_server_destroy": [vp],
        "raja_cdna4_server_start": [vp],
        "raja_cdna4_server_submit": [vp, ctypes.POINTER(Request), i64, ctypes.POINTER(i64)],
        "raja_cdna4_server_poll": [vp, ctypes.POINTER(Completion), i64, dbl, ctypes.POINTER(i64)],
        "raja_cdna4_server_stop": [vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.raja_cdna4_last_error.restype = ctypes.c_char_p
    L.raja_cdna4_last_error.argtypes = []
    L.raja_cdna4_version.restype = ctypes.c_char_p
    L.raja_cdna4_version.argtypes = []


def lib():
    """Load libraja_cdna4.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RajaError("libraja_cdna4.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise RajaError("raja_cdna4 error %d: %s" % (
            status, lib().raja_cdna4_last_error().decode()))


_initialized = set()


def _prep(*tensors):
    """Validate GPU tensors and return the current stream handle."""
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise RajaError("raja_amd operates on GPU tensors only")
        if not t.is_contiguous():
            raise RajaError("spans must be contiguous")
        dev = t.device.index if dev is None else dev
    if dev is None:
        dev = torch.cuda.current_device()
    if dev not in _initialized:
        _check(lib().raja_cdna4_init(dev))
        _initialized.add(dev)
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def device_status():
    """Raise if a scan or radix pass gave up waiting on a lookback descriptor
    since the last check (waits for the device first)."""
    _check(lib().raja_cdna4_device_status())


def _dt(t):
    try:
        return _DTYPES[t.dtype]
    except KeyError:
        raise RajaError("unsupported dtype %s" % t.dtype)


def _same_device(*tensors):
    """Every tensor (None skipped) on one device; the C entry points take raw
    pointers and no lengths beyond n, so this and the size checks below run
    before any call."""
    devs = {t.device for t in tensors if t is not None}
    if len(devs) > 1:
        raise RajaError("tensors on different devices: %s" % sorted(str(d) for d in devs))


def _check_out(out, dtype, numel, what="out"):
    if out.dtype != dtype:
        raise RajaError("%s has dtype %s, expected %s" % (what, out.dtype, dtype))
    if out.numel() < numel:
        raise RajaError("%s holds %d elements, needs at least %d" % (what, out.numel(), numel))


def _check_plan(plan, *tensors):
    """A segment plan's calls read plan.total elements of every span and
    write plan.nseg results; the spans must be on the plan's device."""
    if getattr(plan, "handle", None) is None:
        raise RajaError("segment plan is closed")
    want = torch.device("cuda", plan.device)
    for t in tensors:
        if t is not None and t.device != want:
            raise RajaError("tensor on %s, plan on %s" % (t.device, want))


def _host_scalar(value, dtype):
    buf = torch.tensor([value], dtype=dtype)
    return buf, ctypes.c_void_p(buf.data_ptr())


_IDENT = {
    operators.plus: lambda dt: 0,
    operators.minimum: lambda dt: (torch.finfo(dt).max if dt.is_floating_point
                                   else torch.iinfo(dt).max),
    operators.maximum: lambda dt: (-torch.finfo(dt).max if dt.is_floating_point
                                   else torch.iinfo(dt).min),
}


# ---- scans (include/RAJA/pattern/scan.hpp:49-441) ---------------------------

def inclusive_scan(in_, out, op=operators.plus):
    if in_.numel() != out.numel() or in_.dtype != out.dtype:
        raise RajaError("in and out spans must match in length and type")
    _same_device(in_, out)
    s = _prep(in_, out)
    _check(lib().raja_cdna4_inclusive_scan(_dt(in_), op, in_.data_ptr(), out.data_ptr(),
                                           in_.numel(), s))
    return out


def exclusive_scan(in_, out, op=operators.plus, init=None):
    if in_.numel() != out.numel() or in_.dtype != out.dtype:
        raise RajaError("in and out spans must match in length and type")
    _same_device(in_, out)
    s = _prep(in_, out)
    if init is None:
        init = _IDENT[op](in_.dtype)
    hold, p = _host_scalar(init, in_.dtype)
    _check(lib().raja_cdna4_exclusive_scan(_dt(in_), op, in_.data_ptr(), out.data_ptr(),
                                           in_.numel(), p, s))
    return out


def exclusive_scan_noalloc(in_, out, op=operators.plus, init=None):
    """exclusive_scan through raja_cdna4_exclusive_scan_noalloc: no device
    allocation (a fixed per-GPU scratch array); calls on one GPU must not
    overlap."""
    if in_.numel() != out.numel() or in_.dtype != out.dtype:
        raise RajaError("in and out spans must match in length and type")
    _same_device(in_, out)
    s = _prep(in_, out)
    if init is None:
        init = _IDENT[op](in_.dtype)
    hold, p = _host_scalar(init, in_.dtype)
    _check(lib().raja_cdna4_exclusive_scan_noalloc(_dt(in_), op, in_.data_ptr(), out.data_ptr(),
                                                   in_.numel(), p, s))
    return out


def inclusive_scan_inplace(x, op=operators.plus):
    return inclusive_scan(x, x, op)


def exclusive_scan_inplace(x, op=operators.plus, init=None):
    return exclusive_scan(x, x, op, init)


# ---- sorts (include/RAJA/pattern/sort.hpp:49-453) ------------------------------

def sort_pairs(keys, vals=None, comparator=operators.less):
    if comparator not in (operators.less, operators.greater):
        raise RajaError("GPU sorts accept operators.less or operators.greater")
    if vals is not None and vals.numel() != keys.numel():
        raise RajaError("keys and values must have the same length")
    if vals is not None and vals.element_size() not in (4, 8):
        raise RajaError("values must be 4- or 8-byte elements")
    _same_device(keys, vals)
    s = _prep(keys, vals)
    vb = 0 if vals is None else vals.element_size()
    _check(lib().raja_cdna4_sort_pairs(_dt(keys), keys.data_ptr(),
                                       None if vals is None else vals.data_ptr(), vb,
                                       keys.numel(), 1 if comparator == operators.greater else 0,
                                       s))
    return keys, vals


stable_sort_pairs = sort_pairs  # the LSD radix sort is stable


def sort(keys, comparator=operators.less):
    sort_pairs(keys, None, comparator)
    return keys


stable_sort = sort


# ---- reductions (include/RAJA/pattern/reduce.hpp:79-233) ----------------------

def reduce(x, op, init=None, out=None):
    """Device result op(init, fold_op(x)) as a 1-element tensor."""
    if out is not None:
        _check_out(out, x.dtype, 1)
        _same_device(x, out)
    s = _prep(x, out)
    if init is None:
        init = {operators.bit_or: 0, operators.bit_and: -1}.get(op, None)
        if init is None:
            init = _IDENT[op](x.dtype)
    hold, p = _host_scalar(init, x.dtype)
    if out is None:
        out = torch.empty(1, dtype=x.dtype, device=x.device)
    _check(lib().raja_cdna4_reduce(_dt(x), op, x.data_ptr(), x.numel(), p, out.data_ptr(), s))
    return out


def reduce_sum(x, init=0):
    return reduce(x, operators.plus, init)


def reduce_min(x, init=None):
    return reduce(x, operators.minimum, init)


def reduce_max(x, init=None):
    return reduce(x, operators.maximum, init)


def _reduce_loc(x, is_max, init_val, init_loc):
    s = _prep(x)
    if init_val is None:
        init_val = _IDENT[operators.maximum if is_max else operators.minimum](x.dtype)
    hold, p = _host_scalar(init_val, x.dtype)
    val = torch.empty(1, dtype=x.dtype, device=x.device)
    loc = torch.empty(1, dtype=torch.int64, device=x.device)
    _check(lib().raja_cdna4_reduce_loc(_dt(x), is_max, x.data_ptr(), x.numel(), p, init_loc,
                                       val.data_ptr(), loc.data_ptr(), s))
    return val, loc


def reduce_minloc(x, init_val=None, init_loc=-1):
    return _reduce_loc(x, 0, init_val, init_loc)


def reduce_maxloc(x, init_val=None, init_loc=-1):
    return _reduce_loc(x, 1, init_val, init_loc)


def multi_reduce_f64(x, out=None):
    """(sum, min, max, minloc_val, minloc_loc) of FP64 x in one pass; returns a
    6-slot float64 device tensor (slot 4 holds the int64 loc bits)."""
    if x.dtype != torch.float64:
        raise RajaError("multi_reduce_f64 needs float64")
    if out is not None:
        _check_out(out, torch.float64, 6)
        _same_device(x, out)
    s = _prep(x, out)
    if out is None:
        out = torch.empty(6, dtype=torch.float64, device=x.device)
    _check(lib().raja_cdna4_multi_reduce_f64(x.data_ptr(), x.numel(), out.data_ptr(), s))
    return out


def unpack_multi(out):
    h = out.cpu()
    loc = h[4:5].view(torch.int64).item()
    return {"sum": h[0].item(), "min": h[1].item(), "max": h[2].item(),
            "minloc_val": h[3].item(), "minloc_loc": loc}


# ---- batched (segmented) work ----------------------------------------------------

class Segments:
    """A segment plan: offsets (nseg + 1 host integers) -> device tile tables."""

    def __init__(self, offsets):
        import numpy as np
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if off.ndim != 1 or off.size < 1:
            raise RajaError("offsets must be a 1-D array of nseg + 1 entries")
        self.offsets = off
        self.nseg = off.size - 1
        self.total = int(off[-1])
        h = ctypes.c_void_p()
        _prep()
        self.device = torch.cuda.current_device()
        _check(lib().raja_cdna4_segments_create(off.ctypes.data_as(ctypes.c_void_p), self.nseg,
                                                ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _check(lib().raja_cdna4_segments_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A request server (raja_cdna4_server_*): arena admission, batching of
    small requests, large ones alone on round-robin streams."""

    def __init__(self, mode=SERVE_BATCHED, streams=4, max_inflight=8, max_batch=512,
                 batch_limit=1 << 16, arena_bytes=1 << 30, honor_arrivals=False, batch_bytes=0):
        self.config = ServeConfig(mode, streams, max_inflight, max_batch, batch_limit,
                                  arena_bytes, 1 if honor_arrivals else 0, 0, int(batch_bytes))
        _prep()
        h = ctypes.c_void_p()
        _check(lib().raja_cdna4_server_create(ctypes.byref(self.config), ctypes.byref(h)))
        self.handle = h

    def run(self, requests):
        """requests: a ctypes array of Request (or a list of them).  Blocks
        until all complete; returns (results array, wall seconds)."""
        if not isinstance(requests, ctypes.Array):
            requests = (Request * len(requests))(*requests)
        n = len(requests)
        results = (Result * max(n, 1))()
        wall = ctypes.c_double(0.0)
        torch.cuda.synchronize()  # inputs written by torch on its streams are visible
        _check(lib().raja_cdna4_server_run(self.handle, requests, n, results, ctypes.byref(wall)))
        return results, wall.value

    # ---- open serving: a scheduling thread admits requests as they come ----

    def start(self):
        """Start the server's scheduling thread (raja_cdna4_server_start)."""
        _check(lib().raja_cdna4_server_start(self.handle))

    def submit(self, requests):
        """Queue requests (a Request, a list or a ctypes array); returns their
        tickets.  Their buffers must stay alive until they complete."""
        if isinstance(requests, Request):
            requests = [requests]
        if not isinstance(requests, ctypes.Array):
            requests = (Request * len(requests))(*requests)
        n = len(requests)
        tickets = (ctypes.c_int64 * max(n, 1))()
        _check(lib().raja_cdna4_server_submit(self.handle, requests, n, tickets))
        return list(tickets[:n])

    def poll(self, max_count=1024, timeout_s=0.0):
        """Completions ready now (or within timeout_s): a list of
        (ticket, Result)."""
        out = (Completion * max(max_count, 1))()
        cnt = ctypes.c_int64(0)
        _check(lib().raja_cdna4_server_poll(self.handle, out, max_count, float(timeout_s),
                                             ctypes.byref(cnt)))
        return [(out[i].ticket, out[i].result) for i in range(cnt.value)]

    def stop(self):
        """Wait for every submitted request, then end the scheduling thread."""
        _check(lib().raja_cdna4_server_stop(self.handle))

    def close(self):
        if self.handle:
            _check(lib().raja_cdna4_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request(kind, n, in_, out, in_vals=None, out_vals=None, arrival_s=0.0):
    """One Request from raw device addresses (ints) or tensors."""
    def addr(t):
        if t is None:
            return None
        return t.data_ptr() if isinstance(t, torch.Tensor) else int(t)
    return Request(kind, 0, int(n), addr(in_), addr(in_vals), addr(out), addr(out_vals),
                   float(arrival_s))


def segmented_reduce(plan, x, op=operators.plus, init=None, out=None):
    if x.numel() != plan.total:
        raise RajaError("x holds %d elements, the plan covers %d" % (x.numel(), plan.total))
    if out is not None:
        _check_out(out, x.dtype, plan.nseg)
    _check_plan(plan, x, out)
    s = _prep(x, out)
    if init is None:
        init = _IDENT[op](x.dtype)
    hold, p = _host_scalar(init, x.dtype)
    made = out is None
    if made:
        out = torch.empty(max(plan.nseg, 1), dtype=x.dtype, device=x.device)
    _check(lib().raja_cdna4_segmented_reduce(plan.handle, _dt(x), op, x.data_ptr(), p,
                                             out.data_ptr(), s))
    return out[:plan.nseg] if made else out


def _check_plan_scan(plan, in_, out):
    if in_.numel() != plan.total or out.numel() != plan.total:
        raise RajaError("scan spans hold %d and %d elements, the plan covers %d"
                        % (in_.numel(), out.numel(), plan.total))
    if in_.dtype != out.dtype:
        raise RajaError("in and out spans must have the same dtype")
    _check_plan(plan, in_, out)


def segmented_exclusive_scan(plan, in_, out, op=operators.plus, init=None):
    _check_plan_scan(plan, in_, out)
    s = _prep(in_, out)
    if init is None:
        init = _IDENT[op](in_.dtype)
    hold, p = _host_scalar(init, in_.dtype)
    _check(lib().raja_cdna4_segmented_exclusive_scan(plan.handle, _dt(in_), op, in_.data_ptr(),
                                                     out.data_ptr(), p, s))
    return out


def segmented_inclusive_scan(plan, in_, out, op=operators.plus):
    _check_plan_scan(plan, in_, out)
    s = _prep(in_, out)
    _check(lib().raja_cdna4_segmented_inclusive_scan(plan.handle, _dt(in_), op, in_.data_ptr(),
                                                     out.data_ptr(), s))
    return out


def segmented_sort_pairs(plan, keys, vals=None, comparator=operators.less):
    if comparator not in (operators.less, operators.greater):
        raise RajaError("GPU sorts accept operators.less or operators.greater")
    if keys.numel() != plan.total:
        raise RajaError("keys hold %d elements, the plan covers %d" % (keys.numel(), plan.total))
    if vals is not None and vals.numel() != keys.numel():
        raise RajaError("keys and values must have the same length")
    if vals is not None and vals.element_size() not in (4, 8):
        raise RajaError("values must be 4- or 8-byte elements")
    _check_plan(plan, keys, vals)
    s = _prep(keys, vals)
    vb = 0 if vals is None else vals.element_size()
    _check(lib().raja_cdna4_segmented_sort_pairs(
        plan.handle, _dt(keys), keys.data_ptr(), None if vals is None else vals.data_ptr(), vb,
        1 if comparator == operators.greater else 0, s))
    return keys, vals


# ---- shape tuning (include/raja_cdna4.h: raja_cdna4_autotune) ----------------------

TUNE_INCLUSIVE_SCAN, TUNE_EXCLUSIVE_SCAN, TUNE_SORT_PAIRS = 0, 1, 2


def autotune(op, dtype, n):
    """Measure the candidate shapes of (op, dtype) at length n on the current
    device's current stream and record the fastest in the on-disk cache
    (synchronous)."""
    s = _prep()
    _check(lib().raja_cdna4_autotune(op, _DTYPES[dtype], int(n), s))


def tune_query(op, dtype, n):
    """(choice, tuned): the shape a call of (op, dtype, n) takes; choice 0 is
    the default shape."""
    _prep()
    c, t = ctypes.c_int(), ctypes.c_int()
    _check(lib().raja_cdna4_tune_query(op, _DTYPES[dtype], int(n), ctypes.byref(c), ctypes.byref(t)))
    return c.value, bool(t.value)


def tune_cache_path():
    _prep()
    buf = ctypes.create_string_buffer(4096)
    _check(lib().raja_cdna4_tune_cache_path(buf, len(buf)))
    return buf.value.decode()


def tune_forget():
    _check(lib().raja_cdna4_tune_forget())


# ---- synthetic inputs -------------------------------------------------------------

FILL_U32_BITS, FILL_I32, FILL_I64, FILL_F64_UNIFORM, FILL_F64_NORMAL, FILL_I32_IOTA, \
    FILL_F64_IOTA, FILL_U32_CONST = range(8)


_FILL_BYTES = {0: 4, 1: 4, 2: 8, 3: 8, 4: 8, 5: 4, 6: 8, 7: 4}


def fill(kind, out, seed=0, a=0.0, b=0.0):
    if kind not in _FILL_BYTES:
        raise RajaError("unknown fill kind %r" % (kind,))
    if out.element_size() != _FILL_BYTES[kind]:
        raise RajaError("fill kind %d writes %d-byte elements, out has %d-byte ones"
                        % (kind, _FILL_BYTES[kind], out.element_size()))
    s = _prep(out)
    _check(lib().raja_cdna4_fill(kind, out.data_ptr(), out.numel(), seed, float(a), float(b), s))
    return out


# ---- kernel timing (HIP events on the launch stream) ----------------------------------

def timing_enable(on=True):
    _check(lib().raja_cdna4_timing_enable(1 if on else 0))


def timing_reset():
    _check(lib().raja_cdna4_timing_reset())


def timing_query(name):
    c = ctypes.c_int64()
    t, mn, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib().raja_cdna4_timing_query(name.encode(), ctypes.byref(c), ctypes.byref(t),
                                         ctypes.byref(mn), ctypes.byref(mx)))
    return {"count": c.value, "total_ms": t.value, "min_ms": mn.value, "max_ms": mx.value}
