// Resources: Host and Cdna4 (a HIP stream on an MI355X), Event and
// EventProxy.  Counterpart of the camp resources the reference wraps
// (include/RAJA/util/resource.hpp:38-240): get_default(), get_stream(),
// wait(), wait_for(Event*), get_event(), allocate/deallocate, memcpy,
// memset.  Cdna4 allocations come from libraja_cdna4.so's caching pools.
#pragma once

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "RAJA/config.hpp"

#if defined(RAJA_ENABLE_CDNA4)
#include "raja_cdna4.h"
#endif

namespace RAJA
{
namespace resources
{

enum struct MemoryAccess
{
  Unknown,
  Device,
  Pinned,
  Managed
};

enum struct Platform
{
  undefined = 0,
  host = 1,
  cdna4 = 2
};

class Host
{
public:
  static Host get_default() { return Host(); }
  Platform get_platform() const { return Platform::host; }
  void wait() {}
  template<typename T>
  T* allocate(size_t n, MemoryAccess = MemoryAccess::Device)
  {
    return static_cast<T*>(std::malloc(n * sizeof(T) + (n == 0)));
  }
  void* calloc(size_t bytes, MemoryAccess = MemoryAccess::Device) { return std::calloc(1, bytes + 1); }
  void deallocate(void* p, MemoryAccess = MemoryAccess::Device) { std::free(p); }
  void memcpy(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); }
  void memset(void* p, int v, size_t bytes) { std::memset(p, v, bytes); }
};

#if defined(RAJA_ENABLE_CDNA4)

namespace detail
{
inline void check_hip(hipError_t e, const char* what)
{
  if (e != hipSuccess)
  {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}
inline void check_lib(int status, const char* what)
{
  if (status != RAJA_CDNA4_SUCCESS)
  {
    throw std::runtime_error(std::string(what) + ": " + raja_cdna4_last_error());
  }
}
// A small pool of default-priority streams per process, handed out round
// robin to new Cdna4 resources.
inline hipStream_t pool_stream(int which)
{
  static hipStream_t streams[16] = {};
  static bool made[16] = {};
  const int k = which & 15;
  if (!made[k])
  {
    check_hip(hipStreamCreateWithFlags(&streams[k], hipStreamNonBlocking), "hipStreamCreate");
    made[k] = true;
  }
  return streams[k];
}
inline int next_stream_index()
{
  static int counter = 1;  // index 0 is the default resource's stream
  return counter++ % 16;
}
}  // namespace detail

class Event;

class Cdna4
{
public:
  Cdna4() : m_stream(detail::pool_stream(detail::next_stream_index())) {}
  explicit Cdna4(hipStream_t s) : m_stream(s) {}
  static Cdna4 get_default() { return Cdna4(detail::pool_stream(0)); }
  static Cdna4 CdnaFromStream(hipStream_t s) { return Cdna4(s); }
  Platform get_platform() const { return Platform::cdna4; }
  hipStream_t get_stream() const { return m_stream; }
  void wait() { detail::check_hip(hipStreamSynchronize(m_stream), "hipStreamSynchronize"); }
  inline Event get_event();
  inline void wait_for(Event* e);
  template<typename T>
  T* allocate(size_t n, MemoryAccess ma = MemoryAccess::Device)
  {
    void* p = nullptr;
    if (ma == MemoryAccess::Pinned)
    {
      detail::check_lib(raja_cdna4_pinned_malloc(&p, n * sizeof(T)), "pinned allocate");
    }
    else
    {
      detail::check_lib(raja_cdna4_device_malloc(&p, n * sizeof(T)), "device allocate");
    }
    return static_cast<T*>(p);
  }
  void deallocate(void* p, MemoryAccess ma = MemoryAccess::Device)
  {
    // Pool memory is reused by later allocations: order this after the
    // work already queued on the stream, as hipFree would.
    wait();
    if (ma == MemoryAccess::Pinned)
    {
      detail::check_lib(raja_cdna4_pinned_free(p), "pinned free");
    }
    else
    {
      detail::check_lib(raja_cdna4_device_free(p), "device free");
    }
  }
  void memcpy(void* dst, const void* src, size_t bytes)
  {
    detail::check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, m_stream), "hipMemcpyAsync");
  }
  void memset(void* p, int v, size_t bytes)
  {
    detail::check_hip(hipMemsetAsync(p, v, bytes, m_stream), "hipMemsetAsync");
  }
  bool operator==(const Cdna4& o) const { return m_stream == o.m_stream; }

private:
  hipStream_t m_stream;
};

class Event
{
public:
  Event() = default;
  explicit Event(hipStream_t s)
  {
    detail::check_hip(hipEventCreateWithFlags(&m_event, hipEventDisableTiming), "hipEventCreate");
    detail::check_hip(hipEventRecord(m_event, s), "hipEventRecord");
  }
  bool check() const { return !m_event || hipEventQuery(m_event) == hipSuccess; }
  void wait() const
  {
    if (m_event) detail::check_hip(hipEventSynchronize(m_event), "hipEventSynchronize");
  }
  hipEvent_t get() const { return m_event; }

private:
  hipEvent_t m_event = nullptr;
};

inline Event Cdna4::get_event() { return Event(m_stream); }

inline void Cdna4::wait_for(Event* e)
{
  if (e && e->get())
  {
    detail::check_hip(hipStreamWaitEvent(m_stream, e->get(), 0), "hipStreamWaitEvent");
  }
}

#else

class Event
{
public:
  bool check() const { return true; }
  void wait() const {}
};

#endif

// Result of an algorithm call: the resource it ran on, convertible to an
// Event that completes with the call.
template<typename Res>
class EventProxy
{
public:
  explicit EventProxy(Res r) : m_res(r) {}
  Res get_resource() const { return m_res; }
  Event get_event()
  {
#if defined(RAJA_ENABLE_CDNA4)
    return get_event_impl(m_res);
#else
    return Event();
#endif
  }
  operator Event() { return get_event(); }
  void wait() { m_res.wait(); }

private:
#if defined(RAJA_ENABLE_CDNA4)
  static Event get_event_impl(Host) { return Event(); }
  static Event get_event_impl(Cdna4& r) { return r.get_event(); }
#endif
  Res m_res;
};

}  // namespace resources

// RAJA::Platform, as in the reference (include/RAJA/util/types.hpp).
using resources::Platform;

namespace type_traits
{
template<typename T>
struct is_resource : std::false_type
{
};
template<>
struct is_resource<resources::Host> : std::true_type
{
};
#if defined(RAJA_ENABLE_CDNA4)
template<>
struct is_resource<resources::Cdna4> : std::true_type
{
};
#endif
}  // namespace type_traits

}  // namespace RAJA
