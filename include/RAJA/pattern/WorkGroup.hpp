// WorkPool / WorkGroup / WorkSite: collect many small loops, then run them
// together (include/RAJA/pattern/WorkGroup.hpp:117-530 and
// include/RAJA/policy/WorkGroup.hpp:41-118 in the reference).
//
//   WorkPool<WorkGroupPolicy<Exec, Order, Storage, Dispatch>, Index,
//            xargs<Args...>, Allocator> pool(allocator);
//   pool.enqueue(RAJA::TypedRangeSegment<Index>(b, e), body);  // empty: ignored
//   auto group = pool.instantiate();   // the pool is emptied for reuse
//   auto site  = group.run(args...);   // body(i, args...) for every loop
//
// Semantics kept: `ordered` runs loops in enqueue order, `reverse_ordered`
// in reverse, unordered policies in any order (equal results for
// independent loops); instantiate() moves the loops out and the pool
// reserves its previous maxima; the WorkSite must outlive an async run.
// Loops are stored once, as copies of the bodies in one byte buffer; the
// storage policies are accepted for API compatibility and all use that
// layout.  Segments are contiguous ranges.
#pragma once

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/index/RangeSegment.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/resource.hpp"

namespace camp
{
// Type list used by direct_dispatch<camp::list<Segment, Body>...> and xargs.
template<typename... Ts>
struct list
{
};
}  // namespace camp

namespace RAJA
{

template<typename... Args>
using xargs = camp::list<Args...>;

namespace policy
{
namespace workgroup
{
struct ordered
{
};
struct reverse_ordered
{
};
struct array_of_pointers
{
};
struct ragged_array_of_objects
{
};
struct constant_stride_array_of_objects
{
};
struct indirect_function_call_dispatch
{
};
struct indirect_virtual_function_dispatch
{
};
template<typename... RangeAndCallables>
struct direct_dispatch
{
};

template<typename EXEC_POLICY_T, typename ORDER_POLICY_T, typename STORAGE_POLICY_T,
         typename DISPATCH_POLICY_T = indirect_function_call_dispatch>
struct WorkGroupPolicy
{
  using exec_policy = EXEC_POLICY_T;
  using order_policy = ORDER_POLICY_T;
  using storage_policy = STORAGE_POLICY_T;
  using dispatch_policy = DISPATCH_POLICY_T;
};
}  // namespace workgroup
}  // namespace policy

using policy::workgroup::array_of_pointers;
using policy::workgroup::constant_stride_array_of_objects;
using policy::workgroup::direct_dispatch;
using policy::workgroup::indirect_function_call_dispatch;
using policy::workgroup::indirect_virtual_function_dispatch;
using policy::workgroup::ordered;
using policy::workgroup::ragged_array_of_objects;
using policy::workgroup::reverse_ordered;
using policy::workgroup::WorkGroupPolicy;

namespace detail
{
namespace workgroup
{

// One enqueued loop: its range and where its body lives in the byte buffer.
struct LoopRecord
{
  int64_t begin;
  int64_t len;
  uint64_t body_offset;
  uint32_t type_index;  // position in direct_dispatch's list, or in the type table
  uint32_t pad;
  void* device_fn;      // indirect dispatch: device function pointer (filled at run)
};

template<typename Index, typename... Args>
struct HostCallable
{
  using call_t = void (*)(const void*, Index, Args...);
  using destroy_t = void (*)(void*);
  using device_fn_t = void* (*)();
  call_t call;
  destroy_t destroy;
  device_fn_t device_fn;  // returns the device address of the body's call routine
};

#if defined(RAJA_ENABLE_CDNA4)
// Defined by the cdna4 backend: device address of the call routine of Body.
template<typename Body, typename Index, typename... Args>
void* device_fn_of();
#endif

template<typename Body, typename Index, typename... Args>
void host_call(const void* obj, Index i, Args... args)
{
  (*static_cast<const Body*>(obj))(i, args...);
}

template<typename Body>
void host_destroy(void* obj)
{
  static_cast<Body*>(obj)->~Body();
}

// Loops of one pool/group: body copies in one buffer plus records.
template<typename Index, typename... Args>
struct LoopStorage
{
  std::vector<LoopRecord> loops;
  std::vector<HostCallable<Index, Args...>> host_fns;
  std::vector<const void*> type_keys;  // per loop: identity of its body type
  std::unique_ptr<unsigned char[]> bytes;
  size_t used = 0;
  size_t cap = 0;

  LoopStorage() = default;
  LoopStorage(LoopStorage&&) = default;
  LoopStorage& operator=(LoopStorage&&) = default;
  ~LoopStorage() { clear(); }

  void reserve(size_t nloops, size_t nbytes)
  {
    loops.reserve(nloops);
    host_fns.reserve(nloops);
    type_keys.reserve(nloops);
    grow(nbytes);
  }

  void grow(size_t need)
  {
    if (need <= cap) return;
    size_t ncap = cap ? cap : 4096;
    while (ncap < need) ncap *= 2;
    std::unique_ptr<unsigned char[]> nb(new unsigned char[ncap]);
    // Bodies are relocated bytewise: trivially relocatable for the loop
    // bodies this pattern stores (lambdas capturing values and pointers).
    if (used) std::memcpy(nb.get(), bytes.get(), used);
    bytes.swap(nb);
    cap = ncap;
  }

  // ON_DEVICE: the body runs on the GPU (only the device call routine is
  // instantiated), otherwise on the host.
  template<bool ON_DEVICE, typename Body>
  void add(int64_t begin, int64_t len, uint32_t type_index, Body&& body)
  {
    using B = std::decay_t<Body>;
    const size_t off = (used + 15) & ~size_t(15);
    grow(off + sizeof(B));
    new (bytes.get() + off) B(std::forward<Body>(body));
    used = off + sizeof(B);
    loops.push_back(LoopRecord{begin, len, (uint64_t)off, type_index, 0u, nullptr});
    host_fns.push_back(HostCallable<Index, Args...>{Thunks<ON_DEVICE, B>::host(), &host_destroy<B>,
                                                    Thunks<ON_DEVICE, B>::device()});
    type_keys.push_back(type_key<B>());
  }

  template<bool ON_DEVICE, typename B, typename Enable = void>
  struct Thunks
  {
    static typename HostCallable<Index, Args...>::call_t host() { return &host_call<B, Index, Args...>; }
    static typename HostCallable<Index, Args...>::device_fn_t device() { return nullptr; }
  };
#if defined(RAJA_ENABLE_CDNA4)
  template<typename B>
  struct Thunks<true, B, void>
  {
    static typename HostCallable<Index, Args...>::call_t host() { return nullptr; }
    static typename HostCallable<Index, Args...>::device_fn_t device()
    {
      return &device_fn_of<B, Index, Args...>;
    }
  };
#endif

  template<typename B>
  static const void* type_key()
  {
    static const char k = 0;
    return &k;
  }

  void clear()
  {
    for (size_t l = 0; l < loops.size(); ++l)
    {
      host_fns[l].destroy(bytes.get() + loops[l].body_offset);
    }
    loops.clear();
    host_fns.clear();
    type_keys.clear();
    used = 0;
  }

  size_t size() const { return loops.size(); }
  const void* body(size_t l) const { return bytes.get() + loops[l].body_offset; }
};

template<typename T, typename List>
struct index_in_list;
template<typename T>
struct index_in_list<T, camp::list<>>
{
  static constexpr int value = -1;
};
template<typename T, typename Seg, typename Body, typename... Rest>
struct index_in_list<T, camp::list<camp::list<Seg, Body>, Rest...>>
{
  static constexpr int rest = index_in_list<T, camp::list<Rest...>>::value;
  static constexpr int value = std::is_same<T, Body>::value ? 0 : (rest < 0 ? -1 : rest + 1);
};

template<typename Dispatch>
struct dispatch_list
{
  using type = void;
};
template<typename... RC>
struct dispatch_list<direct_dispatch<RC...>>
{
  using type = camp::list<RC...>;
};

// Records one loop in the storage (the work every runner's enqueue does):
// empty segments are ignored; with direct_dispatch the body's position in
// the dispatch list selects its call at run time.
template<bool ON_DEVICE, typename Dispatch, typename Storage, typename Segment, typename Body>
void enqueue_loop(Storage& storage, Segment&& seg, Body&& body)
{
  using B = std::decay_t<Body>;
  using list_type = typename dispatch_list<Dispatch>::type;
  static_assert(!ON_DEVICE || std::is_trivially_copyable<B>::value,
                "GPU WorkGroup loop bodies are copied bytewise: capture values and pointers");
  const int64_t len = (int64_t)seg.size();
  if (len <= 0) return;
  uint32_t type_index = 0;
  if (!std::is_same<list_type, void>::value)
  {
    constexpr int k = index_in_list<
        B, std::conditional_t<std::is_same<list_type, void>::value, camp::list<>, list_type>>::value;
    static_assert(std::is_same<list_type, void>::value || k >= 0,
                  "direct_dispatch: body type is not in the dispatch list");
    type_index = (uint32_t)(k < 0 ? 0 : k);
  }
  storage.template add<ON_DEVICE>((int64_t)*seg.begin(), len, type_index, std::forward<Body>(body));
}

// Host dispatch: the stored call routine of each loop's body type.
template<typename Index, typename... Args>
struct HostDispatcher
{
  template<typename Storage>
  static void call(const Storage& s, size_t l, Index i, Args... args)
  {
    s.host_fns[l].call(s.body(l), i, args...);
  }
};

}  // namespace workgroup

// The runner of a WorkGroup, specialised per exec policy, with the
// interface of the reference's detail::WorkRunner
// (include/RAJA/pattern/WorkGroup/WorkRunner.hpp:112-330,
// include/RAJA/policy/hip/WorkGroup/WorkRunner.hpp:283-387):
//   enqueue(storage, iterable, body)      record a loop (at WorkPool::enqueue);
//   run(storage, resource, args...) -> per_run_storage
//                                         run every loop (at WorkGroup::run;
//                                         the WorkSite keeps what it returns);
//   clear()                               drop state before reuse or destruction;
// plus dispatcher_type and resource_type.
template<typename EXEC_POLICY_T, typename ORDER_POLICY_T, typename DISPATCH_POLICY_T,
         typename ALLOCATOR_T, typename INDEX_T, typename... Args>
struct WorkRunner;

// Host runners: loops in (reverse) enqueue order, or any order for the
// unordered policies, which run them in enqueue order too.
template<typename ExecPolicy, typename Order, typename Dispatch, typename Alloc, typename Index,
         typename... Args>
struct HostWorkRunner
{
  using resource_type = resources::Host;
  using dispatcher_type = workgroup::HostDispatcher<Index, Args...>;
  using per_run_storage = int;

  template<typename WorkContainer, typename Iterable, typename LoopBody>
  void enqueue(WorkContainer& storage, Iterable&& iter, LoopBody&& loop_body)
  {
    workgroup::enqueue_loop<false, Dispatch>(storage, std::forward<Iterable>(iter),
                                             std::forward<LoopBody>(loop_body));
  }

  template<typename WorkContainer>
  per_run_storage run(WorkContainer const& s, resource_type, Args... args) const
  {
    const size_t n = s.size();
    for (size_t k = 0; k < n; ++k)
    {
      const size_t l = std::is_same<Order, reverse_ordered>::value ? n - 1 - k : k;
      const workgroup::LoopRecord& r = s.loops[l];
      if (std::is_same<ExecPolicy, omp_work>::value)
      {
#if defined(RAJA_ENABLE_OPENMP)
#pragma omp parallel for schedule(static)
#endif
        for (int64_t j = 0; j < r.len; ++j)
        {
          dispatcher_type::call(s, l, Index(r.begin + j), args...);
        }
      }
      else
      {
        for (int64_t j = 0; j < r.len; ++j)
        {
          dispatcher_type::call(s, l, Index(r.begin + j), args...);
        }
      }
    }
    return 0;
  }

  void clear() {}
};

template<typename Order, typename Dispatch, typename Alloc, typename Index, typename... Args>
struct WorkRunner<seq_work, Order, Dispatch, Alloc, Index, Args...>
    : HostWorkRunner<seq_work, Order, Dispatch, Alloc, Index, Args...>
{
};

template<typename Order, typename Dispatch, typename Alloc, typename Index, typename... Args>
struct WorkRunner<omp_work, Order, Dispatch, Alloc, Index, Args...>
    : HostWorkRunner<omp_work, Order, Dispatch, Alloc, Index, Args...>
{
};

}  // namespace detail

template<typename WORKGROUP_POLICY_T, typename INDEX_T, typename EXTRA_ARGS_T,
         typename ALLOCATOR_T>
struct WorkPool;
template<typename WORKGROUP_POLICY_T, typename INDEX_T, typename EXTRA_ARGS_T,
         typename ALLOCATOR_T>
struct WorkGroup;
template<typename WORKGROUP_POLICY_T, typename INDEX_T, typename EXTRA_ARGS_T,
         typename ALLOCATOR_T>
struct WorkSite;

template<typename E, typename O, typename S, typename D, typename INDEX_T, typename... Args,
         typename ALLOCATOR_T>
struct WorkSite<WorkGroupPolicy<E, O, S, D>, INDEX_T, xargs<Args...>, ALLOCATOR_T>
{
  using runner = detail::WorkRunner<E, O, D, ALLOCATOR_T, INDEX_T, Args...>;
  using resource_type = typename runner::resource_type;
  using per_run_storage = typename runner::per_run_storage;
  WorkSite(resource_type r, per_run_storage&& prs) : m_resource(r), m_run_storage(std::move(prs)) {}
  WorkSite(WorkSite&&) = default;
  WorkSite& operator=(WorkSite&&) = default;
  resource_type get_resource() const { return m_resource; }
  void clear() { m_run_storage = per_run_storage(); }

private:
  resource_type m_resource;
  per_run_storage m_run_storage;
};

template<typename E, typename O, typename S, typename D, typename INDEX_T, typename... Args,
         typename ALLOCATOR_T>
struct WorkGroup<WorkGroupPolicy<E, O, S, D>, INDEX_T, xargs<Args...>, ALLOCATOR_T>
{
  using policy = WorkGroupPolicy<E, O, S, D>;
  using index_type = INDEX_T;
  using storage_type = detail::workgroup::LoopStorage<INDEX_T, Args...>;
  using workrunner_type = detail::WorkRunner<E, O, D, ALLOCATOR_T, INDEX_T, Args...>;
  using resource_type = typename workrunner_type::resource_type;
  using worksite_type = WorkSite<policy, INDEX_T, xargs<Args...>, ALLOCATOR_T>;

  WorkGroup(WorkGroup&&) = default;
  WorkGroup& operator=(WorkGroup&&) = default;
  WorkGroup(const WorkGroup&) = delete;
  ~WorkGroup() { m_runner.clear(); }

  worksite_type run(resource_type r, Args... args)
  {
    return worksite_type(r, m_runner.run(m_storage, r, args...));
  }
  worksite_type run(Args... args) { return run(resource_type::get_default(), args...); }

  size_t num_loops() const { return m_storage.size(); }
  void clear()
  {
    m_runner.clear();
    m_storage.clear();
  }

private:
  template<typename, typename, typename, typename>
  friend struct WorkPool;
  WorkGroup(storage_type&& s, workrunner_type&& r) : m_storage(std::move(s)), m_runner(std::move(r)) {}
  storage_type m_storage;
  workrunner_type m_runner;
};

template<typename E, typename O, typename S, typename D, typename INDEX_T, typename... Args,
         typename ALLOCATOR_T>
struct WorkPool<WorkGroupPolicy<E, O, S, D>, INDEX_T, xargs<Args...>, ALLOCATOR_T>
{
  using policy = WorkGroupPolicy<E, O, S, D>;
  using index_type = INDEX_T;
  using workgroup_type = WorkGroup<policy, INDEX_T, xargs<Args...>, ALLOCATOR_T>;
  using storage_type = typename workgroup_type::storage_type;
  using workrunner_type = typename workgroup_type::workrunner_type;

  explicit WorkPool(const ALLOCATOR_T& a = ALLOCATOR_T()) : m_alloc(a) {}
  WorkPool(const WorkPool&) = delete;
  WorkPool(WorkPool&&) = default;

  size_t num_loops() const { return m_storage.size(); }
  size_t storage_bytes() const { return m_storage.used; }
  void reserve(size_t num_loops, size_t storage_bytes)
  {
    m_storage.reserve(num_loops, storage_bytes);
  }

  template<typename Segment, typename Body>
  void enqueue(Segment&& seg, Body&& body)
  {
    if (m_storage.size() == 0) reserve(m_max_loops, m_max_bytes);  // auto-reserve on reuse
    m_runner.enqueue(m_storage, std::forward<Segment>(seg), std::forward<Body>(body));
    m_max_loops = std::max(m_max_loops, m_storage.size());
    m_max_bytes = std::max(m_max_bytes, m_storage.used);
  }

  workgroup_type instantiate()
  {
    workgroup_type g(std::move(m_storage), std::move(m_runner));
    m_storage = storage_type();
    m_runner = workrunner_type();
    return g;
  }

  void clear()
  {
    m_runner.clear();
    m_storage.clear();
  }

private:
  ALLOCATOR_T m_alloc;
  storage_type m_storage;
  workrunner_type m_runner;
  size_t m_max_loops = 0;
  size_t m_max_bytes = 0;
};

}  // namespace RAJA
