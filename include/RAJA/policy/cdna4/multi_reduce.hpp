// Multi (binned) reducers on the cdna4 policies (the role of
// include/RAJA/policy/hip/multi_reduce.hpp in the reference, which combines
// into replicated global tallies with atomics).
//
// Here every launch that captures the reducer (forall or launch on a cdna4
// policy) gets rows of bins in device memory, one row per block, filled
// with the identity on the launch stream before the kernel.  A block adds
// into its own row with workgroup-scope atomics, which the block's XCD L2
// performs locally: no two blocks share an address, so no cross-XCD
// traffic and no contention between blocks.  get() waits for the launch
// streams, reads the rows back and folds them in block order, then launch
// order, into the host totals.  Integer sums, bitwise ops and min/max are
// exact; FP sums depend only on the order of the additions inside a block.
// When one row per block would exceed 256 MiB, blocks share fewer rows
// (row = block % rows) and add with agent-scope atomics instead.
// Contributions to bins outside [0, num_bins) are dropped on the device.
#pragma once

#include <cstring>
#include <mutex>
#include <vector>

#include "RAJA/pattern/multi_reduce.hpp"
#include "RAJA/policy/cdna4/forall.hpp"

namespace RAJA
{
namespace policy
{
namespace cdna4
{
struct cdna4_multi_reduce
{
};
}  // namespace cdna4
}  // namespace policy
using policy::cdna4::cdna4_multi_reduce;
// Name of the reference's default HIP multi-reduce policy family.
using cdna4_multi_reduce_atomic = cdna4_multi_reduce;

namespace cdna4
{
namespace detail
{

template<typename T, typename Op>
__global__ void multi_identity_kernel(T* rows, size_t count)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
  {
    rows[i] = Op::identity();
  }
}

template<int SCOPE, typename T>
__device__ inline void cas_combine_min(T* p, T v)
{
  T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, SCOPE);
  while (v < cur)
  {
    if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, SCOPE)) break;
  }
}

template<int SCOPE, typename T>
__device__ inline void cas_combine_max(T* p, T v)
{
  T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, SCOPE);
  while (cur < v)
  {
    if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, SCOPE)) break;
  }
}

template<int SCOPE, typename T, typename Op>
struct AtomicCombine;
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::sum_op<T>>
{
  __device__ static void apply(T* p, T v) { (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, SCOPE); }
};
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::min_op<T>>
{
  __device__ static void apply(T* p, T v)
  {
    if constexpr (std::is_integral<T>::value)
      (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, SCOPE);
    else
      cas_combine_min<SCOPE>(p, v);
  }
};
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::max_op<T>>
{
  __device__ static void apply(T* p, T v)
  {
    if constexpr (std::is_integral<T>::value)
      (void)__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, SCOPE);
    else
      cas_combine_max<SCOPE>(p, v);
  }
};
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::or_op<T>>
{
  __device__ static void apply(T* p, T v) { (void)__hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, SCOPE); }
};
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::and_op<T>>
{
  __device__ static void apply(T* p, T v) { (void)__hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, SCOPE); }
};

constexpr size_t kMultiRowBudget = (size_t)256 << 20;

}  // namespace detail
}  // namespace cdna4

template<typename T, typename Op>
class Cdna4MultiImpl
{
  struct Launch
  {
    T* rows;
    int nrows;
    hipStream_t stream;
  };
  struct State
  {
    std::vector<T> init;
    std::vector<T> acc;  // completed launches, folded
    std::vector<Launch> launches;
    std::mutex lock;
  };

public:
  using value_type = T;

  template<typename Container>
  explicit Cdna4MultiImpl(const Container& inits) : m_state(new State()), m_owner(true)
  {
    m_state->init.assign(std::begin(inits), std::end(inits));
    m_state->acc.assign(m_state->init.size(), Op::identity());
    m_nbins = m_state->init.size();
  }

  RAJA_HOST_DEVICE Cdna4MultiImpl(const Cdna4MultiImpl& o)
      : m_state(o.m_state), m_rows(o.m_rows), m_nbins(o.m_nbins), m_nrows(o.m_nrows),
        m_shared(o.m_shared), m_owner(false)
  {
#if defined(__HIP_DEVICE_COMPILE__)
    (void)s_fill_kernel;  // the device pass must see the fill kernel to emit it
#else
    RAJA::cdna4::detail::LaunchInfo& li = RAJA::cdna4::detail::launch_info();
    if (li.setup && m_nbins > 0)
    {
      // one row of bins per block, or fewer shared rows beyond the budget
      size_t nrows = (size_t)li.nblocks;
      const size_t row_bytes = m_nbins * sizeof(T);
      if (nrows * row_bytes > RAJA::cdna4::detail::kMultiRowBudget)
      {
        nrows = RAJA::cdna4::detail::kMultiRowBudget / row_bytes;
        if (nrows < 1) nrows = 1;
      }
      const size_t count = nrows * m_nbins;
      void* p = nullptr;
      resources::detail::check_lib(raja_cdna4_device_malloc(&p, count * sizeof(T)), "multi-reduce rows");
      m_rows = static_cast<T*>(p);
      m_nrows = (int)nrows;
      m_shared = nrows < (size_t)li.nblocks;
      fill_identity(m_rows, count, li.stream);
      std::lock_guard<std::mutex> g(m_state->lock);
      m_state->launches.push_back(Launch{m_rows, m_nrows, li.stream});
    }
#endif
  }

  Cdna4MultiImpl& operator=(const Cdna4MultiImpl&) = delete;

  RAJA_HOST_DEVICE ~Cdna4MultiImpl()
  {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (m_owner)
    {
      try
      {
        fold_launches();
      }
      catch (...)
      {
      }
      delete m_state;
    }
#endif
  }

  RAJA_HOST_DEVICE void combine(size_t bin, const T& v) const
  {
#if defined(__HIP_DEVICE_COMPILE__)
    if (!m_rows || bin >= m_nbins) return;
    const int b = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
    if (!m_shared)
    {
      T* p = m_rows + (size_t)b * m_nbins + bin;
      RAJA::cdna4::detail::AtomicCombine<__HIP_MEMORY_SCOPE_WORKGROUP, T, Op>::apply(p, v);
    }
    else
    {
      T* p = m_rows + (size_t)(b % m_nrows) * m_nbins + bin;
      RAJA::cdna4::detail::AtomicCombine<__HIP_MEMORY_SCOPE_AGENT, T, Op>::apply(p, v);
    }
#else
    std::lock_guard<std::mutex> g(m_state->lock);
    m_state->acc[bin] = Op{}(m_state->acc[bin], v);
#endif
  }

  T get(size_t bin) const
  {
    fold_launches();
    return Op{}(m_state->init[bin], m_state->acc[bin]);
  }

  size_t size() const { return m_state->init.size(); }

  template<typename Container>
  void reset(const Container& inits)
  {
    fold_launches();
    m_state->init.assign(std::begin(inits), std::end(inits));
    m_state->acc.assign(m_state->init.size(), Op::identity());
    m_nbins = m_state->init.size();
  }

private:
  static constexpr void (*s_fill_kernel)(T*, size_t) = &RAJA::cdna4::detail::multi_identity_kernel<T, Op>;

  static void fill_identity(T* rows, size_t count, hipStream_t s)
  {
    const size_t want = (count + 255) / 256;
    const unsigned grid = (unsigned)(want < 4096 ? (want > 0 ? want : 1) : 4096);
    hipLaunchKernelGGL((RAJA::cdna4::detail::multi_identity_kernel<T, Op>), dim3(grid), dim3(256), 0, s,
                       rows, count);
    resources::detail::check_hip(hipGetLastError(), "multi-reduce rows fill");
  }

  void fold_launches() const
  {
    std::vector<Launch> ls;
    {
      std::lock_guard<std::mutex> g(m_state->lock);
      ls.swap(m_state->launches);
    }
    std::vector<T> h;
    for (const Launch& L : ls)
    {
      resources::detail::check_hip(hipStreamSynchronize(L.stream), "multi-reduce wait");
      const size_t nb = m_state->acc.size();
      h.resize((size_t)L.nrows * nb);
      resources::detail::check_hip(hipMemcpy(h.data(), L.rows, h.size() * sizeof(T), hipMemcpyDeviceToHost),
                                   "multi-reduce read");
      for (int r = 0; r < L.nrows; ++r)
      {
        for (size_t b = 0; b < nb; ++b) m_state->acc[b] = Op{}(m_state->acc[b], h[(size_t)r * nb + b]);
      }
      resources::detail::check_lib(raja_cdna4_device_free(L.rows), "multi-reduce rows");
    }
  }

  State* m_state = nullptr;
  T* m_rows = nullptr;
  size_t m_nbins = 0;
  int m_nrows = 0;
  bool m_shared = false;
  bool m_owner = false;
};

RAJA_CDNA4_DECLARE_MULTI_REDUCERS(cdna4_multi_reduce, Cdna4MultiImpl)

}  // namespace RAJA
