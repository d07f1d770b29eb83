// Host runtime of libraja_cdna4.so (see runtime.hpp).
#include "runtime.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace rc
{

namespace
{
thread_local char t_last_error[512] = "no error";

std::mutex g_mutex;
std::map<std::pair<int, hipStream_t>, StreamCtx*> g_ctx;
int g_num_cus[64] = {0};

// ---- caching pools ---------------------------------------------------------
// Size classes are powers of two from 256 B.  A freed block goes back to its
// class and is handed out again; memory returns to HIP only at exit.
struct Pool
{
  bool pinned;
  std::mutex m;
  std::unordered_map<void*, size_t> live;          // ptr -> class bytes
  std::map<size_t, std::vector<void*>> free_list;  // class bytes -> blocks

  static size_t size_class(size_t bytes)
  {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }

  int alloc(void** out, size_t bytes)
  {
    const size_t c = size_class(bytes ? bytes : 1);
    {
      std::lock_guard<std::mutex> g(m);
      auto it = free_list.find(c);
      if (it != free_list.end() && !it->second.empty())
      {
        *out = it->second.back();
        it->second.pop_back();
        live[*out] = c;
        return RAJA_CDNA4_SUCCESS;
      }
    }
    void* p = nullptr;
    hipError_t e = pinned ? hipHostMalloc(&p, c, hipHostMallocDefault) : hipMalloc(&p, c);
    if (e != hipSuccess)
    {
      (void)hipGetLastError();
      return set_error(RAJA_CDNA4_ERROR_OUT_OF_MEMORY, "%s of %zu bytes failed: %s",
                       pinned ? "hipHostMalloc" : "hipMalloc", c, hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> g(m);
    live[p] = c;
    *out = p;
    return RAJA_CDNA4_SUCCESS;
  }

  int release(void* p)
  {
    if (!p) return RAJA_CDNA4_SUCCESS;
    std::lock_guard<std::mutex> g(m);
    auto it = live.find(p);
    if (it == live.end())
    {
      return set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "free of a pointer not from this pool");
    }
    free_list[it->second].push_back(p);
    live.erase(it);
    return RAJA_CDNA4_SUCCESS;
  }
};

Pool g_device_pool{false, {}, {}, {}};
Pool g_pinned_pool{true, {}, {}, {}};

// ---- kernel timing -----------------------------------------------------------
struct TimingRecord
{
  int64_t count = 0;
  double total_ms = 0.0;
  double min_ms = 1e300;
  double max_ms = 0.0;
};
struct PendingTiming
{
  std::string name;
  hipEvent_t start, stop;
};
bool g_timing = false;
std::mutex g_timing_mutex;
std::map<std::string, TimingRecord> g_timing_records;
std::vector<PendingTiming> g_pending;

int collect_timings()
{
  std::vector<PendingTiming> pend;
  {
    std::lock_guard<std::mutex> g(g_timing_mutex);
    pend.swap(g_pending);
  }
  for (auto& p : pend)
  {
    RC_HIP(hipEventSynchronize(p.stop));
    float ms = 0.f;
    RC_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    RC_HIP(hipEventDestroy(p.start));
    RC_HIP(hipEventDestroy(p.stop));
    std::lock_guard<std::mutex> g(g_timing_mutex);
    TimingRecord& r = g_timing_records[p.name];
    r.count += 1;
    r.total_ms += ms;
    if (ms < r.min_ms) r.min_ms = ms;
    if (ms > r.max_ms) r.max_ms = ms;
  }
  return RAJA_CDNA4_SUCCESS;
}

int grow(StreamCtx* c, void** buf, size_t* cap, size_t bytes, bool zero)
{
  if (bytes <= *cap) return RAJA_CDNA4_SUCCESS;
  size_t want = align_up(bytes + bytes / 4, 1 << 20);
  if (*buf)
  {
    RC_HIP(hipStreamSynchronize(c->stream));
    RC_HIP(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
  }
  hipError_t e = hipMalloc(buf, want);
  if (e != hipSuccess)
  {
    (void)hipGetLastError();
    *buf = nullptr;
    return set_error(RAJA_CDNA4_ERROR_OUT_OF_MEMORY, "scratch hipMalloc of %zu bytes failed: %s",
                     want, hipGetErrorString(e));
  }
  if (zero)
  {
    RC_HIP(hipMemset(*buf, 0, want));
  }
  *cap = want;
  return RAJA_CDNA4_SUCCESS;
}
}  // namespace

int set_error(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_last_error, sizeof(t_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int hip_error(hipError_t e, const char* what, const char* file, int line)
{
  return set_error(e == hipErrorOutOfMemory ? RAJA_CDNA4_ERROR_OUT_OF_MEMORY
                                            : RAJA_CDNA4_ERROR_HIP,
                   "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
}

int num_cus()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (g_num_cus[dev] == 0)
  {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    {
      n = 256;
    }
    g_num_cus[dev] = n;
  }
  return g_num_cus[dev];
}

int get_ctx(hipStream_t s, StreamCtx** out)
{
  int dev = 0;
  RC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mutex);
  auto key = std::make_pair(dev, s);
  auto it = g_ctx.find(key);
  if (it != g_ctx.end())
  {
    *out = it->second;
    return RAJA_CDNA4_SUCCESS;
  }
  StreamCtx* c = new StreamCtx();
  c->stream = s;
  c->device = dev;
  RC_HIP(hipMalloc(&c->ctl, sizeof(LaunchControl)));
  RC_HIP(hipMemset(c->ctl, 0, sizeof(LaunchControl)));
  RC_HIP(hipHostMalloc(&c->pinned, pinned_staging_bytes, hipHostMallocDefault));
  RC_HIP(hipMalloc(&c->small, small_slot_bytes));
  RC_HIP(hipMemset(c->small, 0, small_slot_bytes));
  g_ctx[key] = c;
  *out = c;
  return RAJA_CDNA4_SUCCESS;
}

int ensure_scan_desc(StreamCtx* c, size_t bytes)
{
  return grow(c, &c->scan_desc, &c->scan_desc_cap, bytes, true);
}

int ensure_sort_desc(StreamCtx* c, size_t bytes)
{
  return grow(c, &c->sort_desc, &c->sort_desc_cap, bytes, true);
}

int ensure_tmp(StreamCtx* c, size_t bytes)
{
  return grow(c, &c->tmp, &c->tmp_cap, bytes, false);
}

bool timing_enabled() { return g_timing; }

TimedScope::TimedScope(const char* n, hipStream_t s) : name(n), stream(s)
{
  if (g_timing && n)  // a null name times nothing (no event is created)
  {
    if (hipEventCreate(&start) == hipSuccess)
    {
      (void)hipEventRecord(start, stream);
    }
    else
    {
      start = nullptr;
    }
  }
}

TimedScope::~TimedScope()
{
  if (start)
  {
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) == hipSuccess && hipEventRecord(stop, stream) == hipSuccess)
    {
      std::lock_guard<std::mutex> g(g_timing_mutex);
      g_pending.push_back(PendingTiming{name, start, stop});
    }
  }
}

namespace
{
struct GateState
{
  std::recursive_mutex m;
  hipEvent_t ev = nullptr;
  hipStream_t last = nullptr;
  bool recorded = false;
};
GateState g_gate[64];

int gate_event(int dev, bool may_create, hipEvent_t* out)
{
  GateState& g = g_gate[dev];
  if (!g.ev && may_create) RC_HIP(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
  *out = g.ev;
  return RAJA_CDNA4_SUCCESS;
}
}  // namespace

int lookback_gate_setup()
{
  int dev = 0;
  RC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return RAJA_CDNA4_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(g_gate[dev].m);
  hipEvent_t ev = nullptr;
  return gate_event(dev, true, &ev);
}

LookbackGate::LookbackGate(hipStream_t s, bool may_create) : stream_(s)
{
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipGetDevice(&device_) != hipSuccess || device_ < 0 || device_ >= 64 ||
      hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
  {
    (void)hipGetLastError();
    return;
  }
  GateState& g = g_gate[device_];
  g.m.lock();
  hipEvent_t ev = nullptr;
  status_ = gate_event(device_, may_create, &ev);
  if (status_ != RAJA_CDNA4_SUCCESS || !ev)
  {
    g.m.unlock();
    return;
  }
  active_ = true;
  if (g.recorded && g.last != s)
  {
    const hipError_t e = hipStreamWaitEvent(s, ev, 0);
    if (e != hipSuccess) status_ = hip_error(e, "hipStreamWaitEvent", __FILE__, __LINE__);
  }
}

LookbackGate::~LookbackGate()
{
  if (!active_) return;
  GateState& g = g_gate[device_];
  if (hipEventRecord(g.ev, stream_) == hipSuccess)
  {
    g.recorded = true;
    g.last = stream_;
  }
  else
  {
    (void)hipGetLastError();
  }
  g.m.unlock();
}

}  // namespace rc

using namespace rc;

extern "C" {

int raja_cdna4_init(int device)
{
  RC_HIP(hipSetDevice(device));
  (void)num_cus();
  StreamCtx* c = nullptr;
  RC_TRY(get_ctx(nullptr, &c));
  RC_TRY(rscan::noalloc_setup(device));
  RC_TRY(lookback_gate_setup());
  return RAJA_CDNA4_SUCCESS;
}

const char* raja_cdna4_last_error(void) { return t_last_error; }

const char* raja_cdna4_version(void) { return "raja-cdna4 0.1 (gfx950)"; }

int raja_cdna4_num_cus(void) { return num_cus(); }

int raja_cdna4_stream_synchronize(void* stream)
{
  RC_HIP(hipStreamSynchronize((hipStream_t)stream));
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_stream_release(void* stream)
{
  int dev = 0;
  RC_HIP(hipGetDevice(&dev));
  StreamCtx* c = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mutex);
    auto it = g_ctx.find(std::make_pair(dev, (hipStream_t)stream));
    if (it == g_ctx.end()) return RAJA_CDNA4_SUCCESS;
    c = it->second;
    g_ctx.erase(it);
  }
  RC_HIP(hipStreamSynchronize(c->stream));
  if (c->ctl) RC_HIP(hipFree(c->ctl));
  if (c->scan_desc) RC_HIP(hipFree(c->scan_desc));
  if (c->sort_desc) RC_HIP(hipFree(c->sort_desc));
  if (c->tmp) RC_HIP(hipFree(c->tmp));
  if (c->pinned) RC_HIP(hipHostFree(c->pinned));
  if (c->small) RC_HIP(hipFree(c->small));
  delete c;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_device_status(void)
{
  RC_HIP(hipDeviceSynchronize());
  uint32_t scans = 0, sorts = 0;
  RC_TRY(rscan::take_spin_timeouts(&scans));
  RC_TRY(rsort::take_spin_timeouts(&sorts));
  if (scans || sorts)
  {
    return set_error(RAJA_CDNA4_ERROR_HIP,
                     "%u scan and %u sort lookback waits gave up after %.1f s: the results of "
                     "those calls are invalid",
                     scans, sorts, (double)RAJA::cdna4::dev::kSpinLimitTicks / 1e8);
  }
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_device_malloc(void** ptr, size_t bytes)
{
  if (!ptr) return set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "ptr is NULL");
  return g_device_pool.alloc(ptr, bytes);
}

int raja_cdna4_device_free(void* ptr) { return g_device_pool.release(ptr); }

int raja_cdna4_pinned_malloc(void** ptr, size_t bytes)
{
  if (!ptr) return set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "ptr is NULL");
  return g_pinned_pool.alloc(ptr, bytes);
}

int raja_cdna4_pinned_free(void* ptr) { return g_pinned_pool.release(ptr); }

int raja_cdna4_timing_enable(int on)
{
  g_timing = (on != 0);
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_timing_reset(void)
{
  RC_TRY(collect_timings());
  std::lock_guard<std::mutex> g(g_timing_mutex);
  g_timing_records.clear();
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_timing_query(const char* name, int64_t* count, double* total_ms,
                            double* min_ms, double* max_ms)
{
  RC_TRY(collect_timings());
  std::lock_guard<std::mutex> g(g_timing_mutex);
  auto it = g_timing_records.find(name ? name : "");
  if (it == g_timing_records.end())
  {
    if (count) *count = 0;
    if (total_ms) *total_ms = 0.0;
    if (min_ms) *min_ms = 0.0;
    if (max_ms) *max_ms = 0.0;
    return RAJA_CDNA4_SUCCESS;
  }
  if (count) *count = it->second.count;
  if (total_ms) *total_ms = it->second.total_ms;
  if (min_ms) *min_ms = it->second.min_ms;
  if (max_ms) *max_ms = it->second.max_ms;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_timing_names(char* buf, size_t len)
{
  RC_TRY(collect_timings());
  std::string s;
  {
    std::lock_guard<std::mutex> g(g_timing_mutex);
    for (auto& kv : g_timing_records)
    {
      if (!s.empty()) s += ",";
      s += kv.first;
    }
  }
  if (!buf || len == 0) return set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "empty buffer");
  snprintf(buf, len, "%s", s.c_str());
  return RAJA_CDNA4_SUCCESS;
}

}  // extern "C"
