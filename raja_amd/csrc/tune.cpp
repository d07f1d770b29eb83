// Shape tuning with an on-disk cache (see tune.hpp).
#include "tune.hpp"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "runtime.hpp"

namespace rtune
{
namespace
{
constexpr int kMaxDevices = 64;
const char* const kOpNames[kNumOps] = {"inclusive_scan", "exclusive_scan", "sort_pairs"};

struct Entry
{
  int choice = 0;
  double ms = 0.0;
};

using Key = std::tuple<int, int, int>;  // op, dtype, bucket

struct DeviceCache
{
  bool loaded = false;
  std::string path;  // empty: nothing is persisted
  std::map<Key, Entry> table;
};

std::mutex g_mutex;
DeviceCache g_dev[kMaxDevices];
std::mutex g_measure_mutex;  // one measurement at a time

bool valid_op(int op) { return op >= 0 && op < kNumOps; }
bool valid_dtype(int d) { return d >= RAJA_CDNA4_I32 && d <= RAJA_CDNA4_F64; }

void candidates(int op, int dtype, int out[3])
{
  if (op == kSortPairs)
    rsort::tile_choices(dtype, out);
  else
    rscan::span_choices(dtype, out);
}

bool is_candidate(int op, int dtype, int c)
{
  int cand[3];
  candidates(op, dtype, cand);
  return c == cand[0] || c == cand[1] || c == cand[2];
}

std::string cache_dir()
{
  if (const char* d = getenv("RAJA_CDNA4_TUNE_DIR")) return d;
  if (const char* x = getenv("XDG_CACHE_HOME")) return std::string(x) + "/raja_cdna4";
  if (const char* h = getenv("HOME")) return std::string(h) + "/.cache/raja_cdna4";
  return "";
}

// The cache file of a device: its architecture, CU count and the HIP
// runtime version are in the name, so another GPU or ROCm never reads it.
int device_path(int dev, std::string* out)
{
  hipDeviceProp_t p;
  RC_HIP(hipGetDeviceProperties(&p, dev));
  int rv = 0;
  RC_HIP(hipRuntimeGetVersion(&rv));
  std::string arch(p.gcnArchName);
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch.resize(colon);
  const std::string dir = cache_dir();
  if (dir.empty())
  {
    out->clear();
    return RAJA_CDNA4_SUCCESS;
  }
  char name[256];
  snprintf(name, sizeof(name), "/tune-%s-%dcu-hip%d.txt", arch.c_str(), p.multiProcessorCount, rv);
  *out = dir + name;
  return RAJA_CDNA4_SUCCESS;
}

int op_from_name(const char* s)
{
  for (int i = 0; i < kNumOps; ++i)
    if (strcmp(s, kOpNames[i]) == 0) return i;
  return -1;
}

// Reads the file once per device (g_mutex held).  Malformed lines, unknown
// operations and shapes that are not candidates are ignored.
int load(int dev)
{
  DeviceCache& c = g_dev[dev];
  if (c.loaded) return RAJA_CDNA4_SUCCESS;
  RC_TRY(device_path(dev, &c.path));
  c.loaded = true;
  if (c.path.empty()) return RAJA_CDNA4_SUCCESS;
  FILE* f = fopen(c.path.c_str(), "r");
  if (!f) return RAJA_CDNA4_SUCCESS;  // no cache yet: default shapes
  char line[512];
  while (fgets(line, sizeof(line), f))
  {
    if (line[0] == '#') continue;
    char op[64];
    int dtype = -1, b = -1, ch = -1;
    double ms = 0.0;
    if (sscanf(line, "%63s %d %d %d %lf", op, &dtype, &b, &ch, &ms) != 5) continue;
    const int o = op_from_name(op);
    if (!valid_op(o) || !valid_dtype(dtype) || b < 10 || b > 40 || !is_candidate(o, dtype, ch))
      continue;
    c.table[Key(o, dtype, b)] = Entry{ch, ms};
  }
  fclose(f);
  return RAJA_CDNA4_SUCCESS;
}

void mkdirs(const std::string& dir)
{
  std::string cur;
  for (size_t i = 0; i <= dir.size(); ++i)
  {
    if (i == dir.size() || dir[i] == '/')
    {
      if (!cur.empty()) (void)mkdir(cur.c_str(), 0755);
    }
    if (i < dir.size()) cur += dir[i];
  }
}

// Writes the device's whole table (g_mutex held): a temporary file renamed
// over the old one, so a reader never sees half a file.
int save(int dev)
{
  DeviceCache& c = g_dev[dev];
  if (c.path.empty()) return RAJA_CDNA4_SUCCESS;
  const size_t slash = c.path.rfind('/');
  if (slash != std::string::npos) mkdirs(c.path.substr(0, slash));
  char tmp[1024];
  snprintf(tmp, sizeof(tmp), "%s.%d.tmp", c.path.c_str(), (int)getpid());
  FILE* f = fopen(tmp, "w");
  if (!f)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "cannot write the shape cache %s: %s",
                         tmp, strerror(errno));
  }
  fprintf(f, "# raja_cdna4 shape cache: operation dtype length-bucket(ceil log2 n) choice ms\n");
  fprintf(f, "# safe to delete; shapes are measured again when asked\n");
  for (const auto& kv : c.table)
  {
    fprintf(f, "%s %d %d %d %.6f\n", kOpNames[std::get<0>(kv.first)], std::get<1>(kv.first),
            std::get<2>(kv.first), kv.second.choice, kv.second.ms);
  }
  const bool ok = fclose(f) == 0;
  if (!ok || rename(tmp, c.path.c_str()) != 0)
  {
    (void)unlink(tmp);
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "cannot replace the shape cache %s",
                         c.path.c_str());
  }
  return RAJA_CDNA4_SUCCESS;
}

int current_device(int* dev)
{
  RC_HIP(hipGetDevice(dev));
  if (*dev < 0 || *dev >= kMaxDevices)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "device %d beyond %d", *dev, kMaxDevices);
  }
  return RAJA_CDNA4_SUCCESS;
}

// Median of `reps` event-timed runs of fn on s (one untimed run first).
template<typename Fn, typename Prep>
int time_runs(hipStream_t s, int reps, Fn fn, Prep prep, double* ms_out)
{
  hipEvent_t a = nullptr, b = nullptr;
  RC_HIP(hipEventCreate(&a));
  RC_HIP(hipEventCreate(&b));
  std::vector<double> t;
  int st = RAJA_CDNA4_SUCCESS;
  for (int r = 0; r <= reps && st == RAJA_CDNA4_SUCCESS; ++r)
  {
    st = prep();
    if (st != RAJA_CDNA4_SUCCESS) break;
    if (hipEventRecord(a, s) != hipSuccess) st = rc::set_error(RAJA_CDNA4_ERROR_HIP, "event record");
    if (st == RAJA_CDNA4_SUCCESS) st = fn();
    if (st == RAJA_CDNA4_SUCCESS && hipEventRecord(b, s) != hipSuccess)
      st = rc::set_error(RAJA_CDNA4_ERROR_HIP, "event record");
    if (st == RAJA_CDNA4_SUCCESS && hipEventSynchronize(b) != hipSuccess)
      st = rc::set_error(RAJA_CDNA4_ERROR_HIP, "event synchronize");
    float ms = 0.f;
    if (st == RAJA_CDNA4_SUCCESS && hipEventElapsedTime(&ms, a, b) != hipSuccess)
      st = rc::set_error(RAJA_CDNA4_ERROR_HIP, "event elapsed time");
    if (st == RAJA_CDNA4_SUCCESS && r > 0) t.push_back(ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st != RAJA_CDNA4_SUCCESS) return st;
  std::sort(t.begin(), t.end());
  *ms_out = t[t.size() / 2];
  return RAJA_CDNA4_SUCCESS;
}

size_t elem_bytes(int dtype)
{
  return (dtype == RAJA_CDNA4_I64 || dtype == RAJA_CDNA4_U64 || dtype == RAJA_CDNA4_F64) ? 8 : 4;
}

// Times every candidate of (op, dtype) on n synthetic elements on stream s
// and returns the fastest; the default shape is kept unless another is at
// least 2 % faster.
int measure(int op, int dtype, int64_t n, hipStream_t s, Entry* best)
{
  int cand[3];
  candidates(op, dtype, cand);
  const size_t eb = elem_bytes(dtype);
  void* a = nullptr;
  void* b = nullptr;
  void* c = nullptr;
  int st = RAJA_CDNA4_SUCCESS;
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipGetLastError();
    st = rc::set_error(e == hipErrorOutOfMemory ? RAJA_CDNA4_ERROR_OUT_OF_MEMORY : RAJA_CDNA4_ERROR_HIP,
                       "shape tuning: %s failed: %s", what, hipGetErrorString(e));
  };
  hipError_t e = hipMalloc(&a, (size_t)n * eb);
  if (e != hipSuccess) fail(e, "hipMalloc");
  if (st == RAJA_CDNA4_SUCCESS && (e = hipMalloc(&b, (size_t)n * 4 > (size_t)n * eb ? (size_t)n * 4 : (size_t)n * eb)) != hipSuccess)
    fail(e, "hipMalloc");
  if (st == RAJA_CDNA4_SUCCESS && op == kSortPairs && (e = hipMalloc(&c, (size_t)n * eb)) != hipSuccess)
    fail(e, "hipMalloc");
  if (st == RAJA_CDNA4_SUCCESS)
  {
    // the sort's keys: uniform 32-bit patterns (fill kind 0); the scans'
    // inputs: integers in [0, 16) (kinds 1 and 2), whatever the element type
    if (op == kSortPairs)
      st = raja_cdna4_fill(0, c, n, 99, 32.0, 0.0, s);
    else
      st = raja_cdna4_fill(eb == 8 ? 2 : 1, a, n, 99, 0.0, 16.0, s);
  }
  Entry out{0, 1e300};
  for (int i = 0; i < 3 && st == RAJA_CDNA4_SUCCESS; ++i)
  {
    if (i > 0 && (cand[i] == 0 || cand[i] == cand[0])) continue;
    double ms = 0.0;
    if (op == kSortPairs)
    {
      st = time_runs(
          s, 5, [&] { return rsort::tune_sort(dtype, a, b, n, cand[i], s); },
          [&]() -> int {
            RC_HIP(hipMemcpyAsync(a, c, (size_t)n * eb, hipMemcpyDeviceToDevice, s));
            return RAJA_CDNA4_SUCCESS;
          },
          &ms);
    }
    else
    {
      st = time_runs(
          s, 5, [&] { return rscan::tune_scan(op == kExclusiveScan, dtype, a, b, n, cand[i], s); },
          [] { return RAJA_CDNA4_SUCCESS; }, &ms);
    }
    if (st != RAJA_CDNA4_SUCCESS) break;
    if (i == 0 || ms < out.ms * 0.98) out = Entry{cand[i], ms};
  }
  if (s) (void)hipStreamSynchronize(s);
  else (void)hipDeviceSynchronize();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (c) (void)hipFree(c);
  if (st == RAJA_CDNA4_SUCCESS) *best = out;
  return st;
}

// Measures (op, dtype) at n and records the result in memory, then in the
// file; the return value reports a failed write, but the choice is kept
// for this process either way.
int tune_now(int dev, int op, int dtype, int64_t n, hipStream_t s, Entry* out)
{
  std::lock_guard<std::mutex> mg(g_measure_mutex);
  Entry e;
  RC_TRY(measure(op, dtype, n, s, &e));
  std::lock_guard<std::mutex> g(g_mutex);
  RC_TRY(load(dev));
  g_dev[dev].table[Key(op, dtype, bucket(n))] = e;
  *out = e;
  return save(dev);
}

bool autotune_env()
{
  const char* v = getenv("RAJA_CDNA4_AUTOTUNE");
  return v && v[0] == '1';
}
}  // namespace

int bucket(int64_t n)
{
  int b = 10;
  while (b < 62 && ((int64_t)1 << b) < n) ++b;
  return b;
}

int choice(int op, int dtype, int64_t n, hipStream_t s, int* out)
{
  *out = 0;
  if (!valid_op(op) || !valid_dtype(dtype) || n <= 0) return RAJA_CDNA4_SUCCESS;
  int dev = 0;
  RC_TRY(current_device(&dev));
  {
    std::lock_guard<std::mutex> g(g_mutex);
    RC_TRY(load(dev));
    auto it = g_dev[dev].table.find(Key(op, dtype, bucket(n)));
    if (it != g_dev[dev].table.end())
    {
      *out = it->second.choice;
      return RAJA_CDNA4_SUCCESS;
    }
  }
  if (!autotune_env()) return RAJA_CDNA4_SUCCESS;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RC_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return RAJA_CDNA4_SUCCESS;
  Entry e;
  const int st = tune_now(dev, op, dtype, n, s, &e);
  if (st != RAJA_CDNA4_SUCCESS && st != RAJA_CDNA4_ERROR_INVALID_ARGUMENT) return st;
  *out = e.choice;  // an unwritable cache file only costs the next process a measurement
  return RAJA_CDNA4_SUCCESS;
}

}  // namespace rtune

using namespace rtune;

extern "C" {

int raja_cdna4_autotune(int tune_op, int dtype, int64_t n, void* stream)
{
  if (!valid_op(tune_op)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "tune op %d", tune_op);
  if (!valid_dtype(dtype)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "dtype %d", dtype);
  if (n <= 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "length %lld", (long long)n);
  if (tune_op == kSortPairs && elem_bytes(dtype) != 4)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "sort shapes are tuned for 4-byte keys");
  }
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RC_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "shapes cannot be measured under capture");
  }
  int dev = 0;
  RC_TRY(current_device(&dev));
  Entry e;
  return tune_now(dev, tune_op, dtype, n, s, &e);
}

int raja_cdna4_tune_query(int tune_op, int dtype, int64_t n, int* choice_out, int* tuned_out)
{
  if (!valid_op(tune_op)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "tune op %d", tune_op);
  if (!valid_dtype(dtype)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "dtype %d", dtype);
  int dev = 0;
  RC_TRY(current_device(&dev));
  std::lock_guard<std::mutex> g(g_mutex);
  RC_TRY(load(dev));
  auto it = g_dev[dev].table.find(Key(tune_op, dtype, bucket(n > 0 ? n : 1)));
  if (choice_out) *choice_out = it == g_dev[dev].table.end() ? 0 : it->second.choice;
  if (tuned_out) *tuned_out = it == g_dev[dev].table.end() ? 0 : 1;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_tune_cache_path(char* buf, size_t len)
{
  if (!buf || len == 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "empty buffer");
  int dev = 0;
  RC_TRY(current_device(&dev));
  std::string p;
  RC_TRY(device_path(dev, &p));
  snprintf(buf, len, "%s", p.c_str());
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_tune_forget(void)
{
  std::lock_guard<std::mutex> g(g_mutex);
  for (auto& d : g_dev)
  {
    d.loaded = false;
    d.path.clear();
    d.table.clear();
  }
  return RAJA_CDNA4_SUCCESS;
}

}  // extern "C"
