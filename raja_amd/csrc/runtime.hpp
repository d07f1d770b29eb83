// Internal host runtime of libraja_cdna4.so: error reporting, per-stream
// scratch, pooled allocation and kernel timing.
//
// Reference counterparts: include/RAJA/policy/hip/MemUtils_HIP.hpp:49-633
// (allocators, launch context, per-stream bookkeeping) and
// include/RAJA/util/basic_mempool.hpp:54-454 (pools).  Design differences:
// scratch is owned per stream, so calls on one stream reuse it in stream
// order without events, and single-pass kernels share one self-resetting
// control block per stream instead of zeroed counters per call.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "RAJA/policy/cdna4/intrinsics.hpp"
#include "raja_cdna4.h"

namespace rc
{

using RAJA::cdna4::dev::LaunchControl;

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_error(hipError_t e, const char* what, const char* file, int line);

#define RC_HIP(expr)                                                           \
  do                                                                           \
  {                                                                            \
    hipError_t rc_e_ = (expr);                                                 \
    if (rc_e_ != hipSuccess) return ::rc::hip_error(rc_e_, #expr, __FILE__, __LINE__); \
  } while (0)

#define RC_TRY(expr)                                                           \
  do                                                                           \
  {                                                                            \
    int rc_s_ = (expr);                                                        \
    if (rc_s_ != RAJA_CDNA4_SUCCESS) return rc_s_;                             \
  } while (0)

// Scratch owned by one (device, stream) pair.  Every call enqueued on that
// stream runs after the previous one finished, so the buffers are reused
// without synchronisation; growth waits for the stream first.
struct StreamCtx
{
  hipStream_t stream = nullptr;
  int device = 0;
  LaunchControl* ctl = nullptr;  // zeroed once; kernels leave it zeroed
  void* scan_desc = nullptr;     // epoch-tagged lookback descriptors (zeroed on growth)
  size_t scan_desc_cap = 0;
  void* sort_desc = nullptr;     // radix lookback descriptors (cleared by kernels)
  size_t sort_desc_cap = 0;
  void* tmp = nullptr;           // general temporaries (alt buffers, partials)
  size_t tmp_cap = 0;
  void* pinned = nullptr;        // 64 KiB pinned host staging
  void* small = nullptr;         // 64 KiB device slots for per-call results
  uint32_t scan_epoch = 0;       // tag of the latest scan's lookback descriptors
};

constexpr size_t pinned_staging_bytes = 64 * 1024;
constexpr size_t small_slot_bytes = 64 * 1024;

int get_ctx(hipStream_t s, StreamCtx** out);
int ensure_scan_desc(StreamCtx* c, size_t bytes);
int ensure_sort_desc(StreamCtx* c, size_t bytes);
int ensure_tmp(StreamCtx* c, size_t bytes);

int num_cus();

// Kernel timing (HIP events around launches on the launch stream).
bool timing_enabled();
struct TimedScope
{
  const char* name;
  hipStream_t stream;
  hipEvent_t start = nullptr;
  TimedScope(const char* n, hipStream_t s);
  ~TimedScope();
};

// One lookback launch at a time per device.  Blocks of a single-pass scan
// or radix pass wait on lower tiles of the same launch; each XCD dispatches
// its share of a launch in order, so one such launch always progresses.
// Two running at once on different streams can deadlock: XCD x fills with
// later blocks of launch A waiting on A's lowest unfinished block, which
// XCD y cannot place because it is full of later blocks of launch B waiting
// on B's lowest block, due on XCD x.  A LookbackGate serialises such work
// across streams: it waits (on the device) for the device's previous gated
// work when that was on another stream, and records an event after its
// own.  Under stream capture it does nothing (a graph replays on one
// stream).  `may_create` false: never creates the device's event (the
// allocation-free entry point), so the gate is skipped until
// raja_cdna4_init or a gated call created it.
class LookbackGate
{
 public:
  LookbackGate(hipStream_t s, bool may_create = true);
  ~LookbackGate();
  int status() const { return status_; }
  LookbackGate(const LookbackGate&) = delete;
  LookbackGate& operator=(const LookbackGate&) = delete;

 private:
  hipStream_t stream_;
  int device_ = -1;
  bool active_ = false;
  int status_ = RAJA_CDNA4_SUCCESS;
};
// Creates the current device's gate event (raja_cdna4_init).
int lookback_gate_setup();

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace rc

// Counts of lookback waits that gave up (SpinBound) since the last call, per
// kernel module; each call reads and clears its module's device counter.
namespace rscan
{
int take_spin_timeouts(uint32_t* count);
// Looks up the current device's copy of the allocation-free scan's fixed
// scratch (once per device).
int noalloc_setup(int device);
}
namespace rsort
{
int take_spin_timeouts(uint32_t* count);
}
