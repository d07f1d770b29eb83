// Multi (binned) reducers on the cdna4 policies (the role of
// include/RAJA/policy/hip/multi_reduce.hpp in the reference, which combines
// into replicated global tallies with atomics, so its FP results vary from
// run to run).
//
// Every launch that captures the reducer (forall or launch on a cdna4
// policy) gets rows of bins in device memory, filled with the identity on
// the launch stream before the kernel:
//   * floating-point types: one row per wave of the launch, written without
//     atomics.  In each combine call the wave groups its active lanes by bin,
//     folds each group's values in lane order, and the group's first lane
//     updates the wave's row (plain load and store: no other wave touches the
//     row, and a wave's calls follow its program order).  The result depends
//     only on the range and the launch shape (block size, grid), never on
//     scheduling: repeats are bit-identical.  Rows for all waves must fit in
//     kMultiRowBudget (1 GiB); beyond it the launch is refused with a message
//     naming a grid that fits (cdna4_exec_grid<B, G>, or fewer teams);
//   * integer types: one row per block, combined with workgroup-scope
//     atomics, which the block's XCD L2 performs locally (integer sums,
//     min/max and bitwise ops give the same value in any order).  When one
//     row per block would exceed the budget, blocks share fewer rows
//     (row = block % rows) and use agent-scope atomics.
// get() waits for the launch streams, folds each launch's rows in row order
// on the device (one thread per bin), reads the folded bins back and
// combines them into the host totals in launch order.  Contributions to bins
// outside [0, num_bins) are dropped on the device.
#pragma once

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
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
  __device__ static void apply(T* p, T v) { (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, SCOPE); }
};
template<int SCOPE, typename T>
struct AtomicCombine<SCOPE, T, reduce::detail::max_op<T>>
{
  __device__ static void apply(T* p, T v) { (void)__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, SCOPE); }
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

constexpr size_t kMultiRowBudget = (size_t)1 << 30;

// out[b] = fold of rows[0..nrows)[b] in row order (one thread per bin).
template<typename T, typename Op>
__global__ void multi_fold_rows_kernel(const T* __restrict__ rows, int nrows, size_t nbins,
                                       T* __restrict__ out)
{
  for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbins;
       b += (size_t)gridDim.x * blockDim.x)
  {
    T r = Op::identity();
    for (int k = 0; k < nrows; ++k) r = Op{}(r, rows[(size_t)k * nbins + b]);
    out[b] = r;
  }
}

// One combine call of a wave into the wave's private row: lanes with the
// same bin fold their values in lane order, the group's first lane updates
// the row.  Every variable the loops test is wave-uniform.
template<typename T, typename Op>
__device__ inline void wave_row_combine(T* row, size_t nbins, size_t bin, T v)
{
  using namespace RAJA::cdna4::dev;
  uint64_t todo = ballot(bin < nbins);
  const int lane = lane_id();
  while (todo)
  {
    const int leader = __builtin_ctzll(todo);
    const size_t lb = read_lane(bin, leader);
    const uint64_t grp = ballot(bin == lb) & todo;
    T r = Op::identity();
    for (uint64_t g = grp; g; g &= g - 1)
    {
      r = Op{}(r, read_lane(v, __builtin_ctzll(g)));
    }
    if (lane == leader) row[lb] = Op{}(row[lb], r);
    todo &= ~grp;
  }
}

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
    (void)s_fill_kernel;  // the device pass must see the fill and fold kernels to emit them
    (void)s_fold_kernel;
#else
    RAJA::cdna4::detail::LaunchInfo& li = RAJA::cdna4::detail::launch_info();
    if (li.probe)
    {
      ++li.probed;
    }
    else if (li.setup && m_nbins > 0)
    {
      const size_t row_bytes = m_nbins * sizeof(T);
      size_t nrows = (size_t)li.nblocks;
      if (kWaveRows)
      {
        // one row of bins per wave of the launch
        const size_t waves_per_block = (size_t)(li.block_threads + 63) / 64;
        nrows = (size_t)li.nblocks * waves_per_block;
        if (nrows * row_bytes > RAJA::cdna4::detail::kMultiRowBudget)
        {
          const size_t fit = RAJA::cdna4::detail::kMultiRowBudget / row_bytes / waves_per_block;
          throw std::runtime_error(
              "cdna4 floating-point multi-reduction: " + std::to_string(m_nbins) + " bins x " +
              std::to_string(nrows) + " waves exceed the " +
              std::to_string(RAJA::cdna4::detail::kMultiRowBudget >> 20) +
              " MiB of per-wave rows; launch at most " + std::to_string(fit) +
              " blocks (cdna4_exec_grid<B, G>, or fewer teams)");
        }
      }
      else if (nrows * row_bytes > RAJA::cdna4::detail::kMultiRowBudget)
      {
        // integers: fewer rows, shared by blocks
        nrows = RAJA::cdna4::detail::kMultiRowBudget / row_bytes;
        if (nrows < 1) nrows = 1;
      }
      const size_t count = nrows * m_nbins;
      void* p = nullptr;
      resources::detail::check_lib(raja_cdna4_device_malloc(&p, count * sizeof(T)), "multi-reduce rows");
      m_rows = static_cast<T*>(p);
      m_nrows = (int)nrows;
      m_shared = !kWaveRows && nrows < (size_t)li.nblocks;
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
    if (!m_rows) return;
    const int b = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
    if constexpr (kWaveRows)
    {
      const int nthreads = (int)(blockDim.x * blockDim.y * blockDim.z);
      const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
      const size_t wave = (size_t)b * (size_t)((nthreads + 63) / 64) + (size_t)(t / 64);
      RAJA::cdna4::detail::wave_row_combine<T, Op>(m_rows + wave * m_nbins, m_nbins, bin, v);
    }
    else
    {
      // integers: the atomic combine gives the same value in any order
      if (bin >= m_nbins) return;
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
  // Floating-point types take per-wave rows without atomics (deterministic).
  static constexpr bool kWaveRows = std::is_floating_point<T>::value;
  static constexpr void (*s_fill_kernel)(T*, size_t) = &RAJA::cdna4::detail::multi_identity_kernel<T, Op>;
  static constexpr void (*s_fold_kernel)(const T*, int, size_t, T*) =
      &RAJA::cdna4::detail::multi_fold_rows_kernel<T, Op>;

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
      const size_t nb = m_state->acc.size();
      // fold the rows in row order on the device into nb values, then read those
      void* folded = nullptr;
      resources::detail::check_lib(raja_cdna4_device_malloc(&folded, nb * sizeof(T)), "multi-reduce fold");
      const size_t want = (nb + 255) / 256;
      const unsigned grid = (unsigned)(want < 4096 ? (want > 0 ? want : 1) : 4096);
      hipLaunchKernelGGL((RAJA::cdna4::detail::multi_fold_rows_kernel<T, Op>), dim3(grid), dim3(256), 0,
                         L.stream, (const T*)L.rows, L.nrows, nb, (T*)folded);
      resources::detail::check_hip(hipGetLastError(), "multi-reduce fold launch");
      h.resize(nb);
      resources::detail::check_hip(hipMemcpyAsync(h.data(), folded, nb * sizeof(T), hipMemcpyDeviceToHost,
                                                  L.stream),
                                   "multi-reduce read");
      resources::detail::check_hip(hipStreamSynchronize(L.stream), "multi-reduce wait");
      for (size_t b = 0; b < nb; ++b) m_state->acc[b] = Op{}(m_state->acc[b], h[b]);
      resources::detail::check_lib(raja_cdna4_device_free(folded), "multi-reduce fold");
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
