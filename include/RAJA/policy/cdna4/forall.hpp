// cdna4 policy family: forall and reducers on an MI355X (gfx950).
//
// Reference counterparts: include/RAJA/policy/hip/policy.hpp:346-359,
// 1246-1358 (hip_exec policies), include/RAJA/policy/hip/forall.hpp:365-639
// (forall kernel and forall_impl), include/RAJA/policy/hip/reduce.hpp:
// 960-1370 (reducers) and include/RAJA/policy/hip/params/reduce.hpp:19-60
// (expt parameter reductions).
//
// Design:
//   * one kernel per loop body: blocks take tiles of 6*B consecutive
//     indices, grid-stride, and a thread handles indices t, t+B, ..., t+5B
//     of its tile, so each wave has six loads in flight on adjacent
//     512-byte segments; the grid is 1024 threads per CU (or fixed by
//     cdna4_exec_grid), never larger than the range needs, so a reducer's
//     result depends on the range, the block size and the CU count only
//     (not on how many registers the body takes);
//   * the kernel makes a private copy of the body; reducers captured in it
//     accumulate in registers and, when that copy is destroyed at kernel
//     end, fold the block's threads with a fixed tree (DPP inside waves, LDS
//     across waves) and store one value per block in pinned host memory;
//   * get() (or, for expt parameters, the end of forall) waits for the
//     stream and folds the block values on the host in block order: the
//     result depends only on the range and the launch shape; no device
//     atomics, no fences;
//   * MinLoc/MaxLoc folds break value ties by the lower index, which is
//     seq_exec's first-index result for increasing iteration order.
#pragma once

#include <algorithm>
#include <mutex>
#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/index/IndexSet.hpp"
#include "RAJA/index/RangeSegment.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/pattern/reduce.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/policy/cdna4/intrinsics.hpp"
#include "RAJA/util/resource.hpp"
#include "raja_cdna4.h"

namespace RAJA
{
namespace policy
{
namespace cdna4
{

template<int BLOCK_SIZE, bool ASYNC = false, int GRID_SIZE = 0>
struct cdna4_exec_t
{
  static_assert(BLOCK_SIZE % 64 == 0 && BLOCK_SIZE <= 1024, "block size: multiple of 64, <= 1024");
  static constexpr Policy policy = Policy::cdna4;
  using resource_type = resources::Cdna4;
  static constexpr int block_size = BLOCK_SIZE;
  static constexpr bool async = ASYNC;
  static constexpr int grid_size = GRID_SIZE;
};

struct cdna4_reduce
{
};

struct cdna4_synchronize
{
};

}  // namespace cdna4
}  // namespace policy

template<int B>
using cdna4_exec = policy::cdna4::cdna4_exec_t<B, false, 0>;
template<int B>
using cdna4_exec_async = policy::cdna4::cdna4_exec_t<B, true, 0>;
template<int B, int G, bool Async = false>
using cdna4_exec_grid = policy::cdna4::cdna4_exec_t<B, Async, G>;
template<int B>
using cdna4_exec_occ_calc = policy::cdna4::cdna4_exec_t<B, false, 0>;
using policy::cdna4::cdna4_reduce;
using policy::cdna4::cdna4_synchronize;

namespace cdna4
{
namespace detail
{

// Launch context for reducer registration: while `setup` is set, a host-side
// copy of a reducer is the copy that travels to the kernel and gets block
// slots; while `probe` is set, a copy only counts itself in `probed` (how a
// launch learns whether a loop body carries reducers before choosing its
// grid).
struct LaunchInfo
{
  bool setup = false;
  int nblocks = 0;
  hipStream_t stream = nullptr;
  int block_threads = 0;  // threads per block of the launch
  bool probe = false;
  int probed = 0;
};

inline LaunchInfo& launch_info()
{
  static thread_local LaunchInfo info;
  return info;
}

// Parallel fold operator of a reducer operator (ValueLoc ties go to the
// lower index).
template<typename Op>
using tree_of = reduce::detail::tree_of<Op>;

// Block fold at kernel end; the block's first thread stores the block value
// at the block's flat index.  Every thread of the block must call it (it
// synchronises the block).  Threads and blocks are numbered x fastest, then
// y, then z.  Full waves fold with the DPP wave tree; a partial last wave
// (block sizes that are not a multiple of 64) folds its lanes in order
// through LDS; the wave values are folded in wave order.
template<typename T, typename TreeOp>
__device__ inline void store_block_value(T v, T* slots)
{
  using namespace RAJA::cdna4::dev;
  static_assert(sizeof(T) % 4 == 0, "cdna4 reducers need 4-byte multiple types");
  // raw LDS bytes: __shared__ objects cannot have constructors
  __shared__ __attribute__((aligned(16))) unsigned char s_raw[sizeof(T) * max_waves];
  __shared__ __attribute__((aligned(16))) unsigned char s_part[sizeof(T) * wave_size];
  T* s_wave = reinterpret_cast<T*>(s_raw);
  T* s_lane = reinterpret_cast<T*>(s_part);
  const TreeOp op{};
  const int nthreads = (int)(blockDim.x * blockDim.y * blockDim.z);
  const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
  const int w = t / wave_size;
  const int nw = (nthreads + wave_size - 1) / wave_size;
  const int last = nthreads - (nw - 1) * wave_size;  // lanes in the last wave
  if (w < nw - 1 || last == wave_size)
  {
    const T r = wave_reduce(v, op);
    if (lane_id() == 0) s_wave[w] = r;
  }
  else
  {
    s_lane[t - w * wave_size] = v;
  }
  __syncthreads();
  if (t == 0)
  {
    if (last != wave_size)
    {
      T r = s_lane[0];
      for (int k = 1; k < last; ++k) r = op(r, s_lane[k]);
      s_wave[nw - 1] = r;
    }
    T r = s_wave[0];
    for (int k = 1; k < nw; ++k) r = op(r, s_wave[k]);
    const int b = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
    slots[b] = r;
  }
  __syncthreads();
}

inline void check(int status, const char* what)
{
  resources::detail::check_lib(status, what);
}

template<typename T>
T* acquire_slots(int n)
{
  void* p = nullptr;
  check(raja_cdna4_pinned_malloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)), "reducer slots");
  return static_cast<T*>(p);
}

inline void release_slots(void* p)
{
  if (p) check(raja_cdna4_pinned_free(p), "reducer slots");
}

}  // namespace detail

// Storage of an old-style reducer for cdna4_reduce.
template<typename T, typename Op>
class Cdna4ReduceImpl
{
  using Tree = typename detail::tree_of<Op>::type;
  struct Launch
  {
    T* slots;
    int nblocks;
    hipStream_t stream;
  };
  struct State
  {
    T value;  // init folded with every completed launch
    T identity;
    std::vector<Launch> launches;
    std::mutex lock;
  };

public:
  using value_type = T;

  Cdna4ReduceImpl(T init, T identity)
      : m_state(new State()), m_identity(identity), m_val(identity), m_owner(true)
  {
    m_state->value = init;
    m_state->identity = identity;
  }

  RAJA_HOST_DEVICE Cdna4ReduceImpl(const Cdna4ReduceImpl& o)
      : m_state(o.m_state), m_slots(o.m_slots), m_identity(o.m_identity), m_val(o.m_identity),
        m_owner(false)
  {
#if defined(__HIP_DEVICE_COMPILE__)
    m_device_private = true;
#else
    detail::LaunchInfo& li = detail::launch_info();
    if (li.probe)
    {
      ++li.probed;
    }
    else if (li.setup)
    {
      m_slots = detail::acquire_slots<T>(li.nblocks);
      std::lock_guard<std::mutex> g(m_state->lock);
      m_state->launches.push_back(Launch{m_slots, li.nblocks, li.stream});
    }
#endif
  }

  Cdna4ReduceImpl& operator=(const Cdna4ReduceImpl&) = delete;

  RAJA_HOST_DEVICE ~Cdna4ReduceImpl()
  {
#if defined(__HIP_DEVICE_COMPILE__)
    if (m_device_private && m_slots)
    {
      detail::store_block_value<T, Tree>(m_val, m_slots);
    }
#else
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

  RAJA_HOST_DEVICE void combine(const T& v) const
  {
#if defined(__HIP_DEVICE_COMPILE__)
    m_val = Op{}(m_val, v);
#else
    m_state->value = Op{}(m_state->value, v);
#endif
  }

  T get() const
  {
    fold_launches();
    return m_state->value;
  }

  void reset(T init, T identity)
  {
    fold_launches();
    m_state->value = init;
    m_state->identity = identity;
    m_identity = identity;
  }

private:
  // Fold each launch's block values in block order, then into the running
  // value with the reducer operator (the running value wins ties).
  void fold_launches() const
  {
    std::vector<Launch> pending;
    {
      std::lock_guard<std::mutex> g(m_state->lock);
      pending.swap(m_state->launches);
    }
    for (const Launch& l : pending)
    {
      resources::detail::check_hip(hipStreamSynchronize(l.stream), "reducer wait");
      T r = m_state->identity;
      for (int b = 0; b < l.nblocks; ++b) r = Tree{}(r, l.slots[b]);
      if (reduce::detail::differs(r, m_state->identity))
      {
        m_state->value = Op{}(m_state->value, r);
      }
      detail::release_slots(l.slots);
    }
  }

  State* m_state = nullptr;
  T* m_slots = nullptr;
  T m_identity;
  mutable T m_val;
  bool m_owner = false;
  bool m_device_private = false;
};

namespace detail
{

// ---- the forall kernel -------------------------------------------------------

// Launch shapes (indices per thread per tile, threads per CU of the default
// grid).  A loop whose body carries reducers (expt::Reduce parameters or
// captured old-style reducers) takes 8 indices per thread at 512 threads per
// CU; any other loop 6 at 1024.  The grid then depends only on the range,
// the block size, the CU count and whether the loop reduces -- never on the
// body's register use -- so both reducer styles over the same loop give the
// same bits.  Measured with B = 256 (indices per thread / threads per CU,
// one MI355X): one forall with four reducers over 2^28 FP64 0.399 ms at
// 4/2048, 0.376 at 6/1024, 0.359 at 8/512, 0.366 at 12/512, 0.40 at
// 16/512; dot + MinLoc over 2^28 0.715 and 0.696 ms at 6/1024 and 8/512;
// daxpy over 2^28 1.206 and 1.362 ms, over 2^24 0.071 and 0.103 ms (two
// consecutive indices per lane, so a read-only body's loads merge into
// 16-byte ones, did not beat 8/512 for reducing loops: 0.361 ms).
constexpr int kForallItems = 6;
constexpr int kForallThreadsPerCU = 1024;
constexpr int kReduceItems = 8;
constexpr int kReduceThreadsPerCU = 512;

inline Index_type forall_grid_cap(int block, int threads_per_cu)
{
  return (Index_type)raja_cdna4_num_cus() * (Index_type)std::max(1, threads_per_cu / block);
}

template<typename EXEC_POL, int BLOCK, int U, typename Seg, typename Body, typename ForallParam>
__global__ __launch_bounds__(BLOCK) void forall_kernel(Seg seg, Body body, ForallParam f_params,
                                                       Index_type len)
{
  Body b(body);  // device-private copy: its reducers fold per block at scope end
  const Index_type tstride = (Index_type)gridDim.x * BLOCK * U;
  for (Index_type base = (Index_type)blockIdx.x * BLOCK * U; base < len; base += tstride)
  {
    if (base + BLOCK * U <= len)
    {
#pragma unroll
      for (int u = 0; u < U; ++u) RAJA::expt::invoke_body(f_params, b, seg.at(base + u * BLOCK + threadIdx.x));
    }
    else
    {
      for (int u = 0; u < U; ++u)
      {
        const Index_type k = base + u * BLOCK + threadIdx.x;
        if (k < len) RAJA::expt::invoke_body(f_params, b, seg.at(k));
      }
    }
  }
  if constexpr (!RAJA::expt::type_traits::is_ForallParamPack_empty<ForallParam>::value)
  {
    RAJA::expt::ParamMultiplexer::parampack_combine(EXEC_POL{}, f_params);
  }
}

template<typename Kernel>
int occupancy_grid(Kernel k, int block)
{
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess ||
      per_cu <= 0)
  {
    (void)hipGetLastError();
    per_cu = 2048 / block;
  }
  return per_cu * raja_cdna4_num_cus();
}

}  // namespace detail
}  // namespace cdna4

RAJA_CDNA4_DECLARE_REDUCERS(cdna4_reduce, cdna4::Cdna4ReduceImpl)

namespace cdna4
{
namespace detail
{
// Launch information handed to param_init / param_resolve (the reference's
// hipInfo, include/RAJA/policy/hip/MemUtils_HIP.hpp): the grid of the launch
// whose reducers are being set up and the resource it runs on.
struct cdna4Info
{
  cdna4Info(resources::Cdna4 r, dim3 grid, dim3 block) : gridDim(grid), blockDim(block), res(r) {}
  dim3 gridDim;
  dim3 blockDim;
  resources::Cdna4 res;
};
}  // namespace detail
}  // namespace cdna4

namespace expt
{
namespace detail
{
// ---- expt parameter reductions on cdna4 policies -----------------------------
// Counterpart of include/RAJA/policy/hip/params/reduce.hpp:19-60: each
// block of the launch folds its threads' ValOps with the fixed block tree
// and stores one value (param_combine, device side, at kernel end); after
// the launch the host waits and folds the block values in block order, then
// combines the result into the target (param_resolve).  No device atomics:
// the result depends only on the range and the launch shape.

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<RAJA::type_traits::is_cdna4_policy<EXEC_POL>::value> param_init(
    EXEC_POL const&, Reducer<OP, T, VOp>& red, const RAJA::cdna4::detail::cdna4Info& info)
{
  const int n = (int)(info.gridDim.x * info.gridDim.y * info.gridDim.z);
  red.block_values = RAJA::cdna4::detail::acquire_slots<T>(n);
  red.nblocks = n;
  red.m_valop = VOp{};
}

template<typename EXEC_POL, typename OP, typename T, typename VOp>
RAJA_HOST_DEVICE std::enable_if_t<RAJA::type_traits::is_cdna4_policy<EXEC_POL>::value> param_combine(
    EXEC_POL const&, Reducer<OP, T, VOp>& red)
{
#if defined(__HIP_DEVICE_COMPILE__)
  RAJA::cdna4::detail::store_block_value<T, reducer_tree<Reducer<OP, T, VOp>>>(red.m_valop.val,
                                                                               red.block_values);
#else
  (void)red;
#endif
}

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<RAJA::type_traits::is_cdna4_policy<EXEC_POL>::value> param_resolve(
    EXEC_POL const&, Reducer<OP, T, VOp>& red, const RAJA::cdna4::detail::cdna4Info& info)
{
  resources::Cdna4 res = info.res;
  res.wait();
  T r = VOp{}.val;
  for (int b = 0; b < red.nblocks; ++b) r = reducer_tree<Reducer<OP, T, VOp>>{}(r, red.block_values[b]);
  red.combineTarget(r);
  RAJA::cdna4::detail::release_slots(red.block_values);
  red.block_values = nullptr;
  red.nblocks = 0;
}
}  // namespace detail
}  // namespace expt

namespace policy
{
namespace cdna4
{

}  // namespace cdna4
}  // namespace policy

namespace detail
{
template<int B, bool A, int G>
struct segit_exec<policy::cdna4::cdna4_exec_t<B, A, G>>
{
  using type = policy::cdna4::cdna4_exec_t<B, true, G>;
  static constexpr bool wait_after = !A;
};
}  // namespace detail

namespace policy
{
namespace cdna4
{

// One forall launch of shape (U indices per thread, TPC threads per CU).
template<int U, int TPC, int B, bool Async, int G, typename Seg, typename BodyT, typename Body,
         typename ForallParam>
void launch_forall(resources::Cdna4& res, const cdna4_exec_t<B, Async, G>& pol, const Seg& seg,
                   Body&& body, ForallParam& f_params, Index_type len)
{
  using EXEC_POL = cdna4_exec_t<B, Async, G>;
  constexpr bool has_params = !expt::type_traits::is_ForallParamPack_empty<ForallParam>::value;
  auto kernel = RAJA::cdna4::detail::forall_kernel<EXEC_POL, B, U, Seg, BodyT, ForallParam>;
  constexpr Index_type per_block = (Index_type)B * U;
  Index_type grid = (len + per_block - 1) / per_block;
  const Index_type cap = (G > 0) ? G : RAJA::cdna4::detail::forall_grid_cap(B, TPC);
  if (grid > cap) grid = cap;
  RAJA::cdna4::detail::cdna4Info info(res, dim3((unsigned)grid), dim3((unsigned)B));
  if constexpr (has_params) expt::ParamMultiplexer::parampack_init(pol, f_params, info);
  hipStream_t stream = res.get_stream();
  RAJA::cdna4::detail::LaunchInfo& li = RAJA::cdna4::detail::launch_info();
  li = RAJA::cdna4::detail::LaunchInfo{true, (int)grid, stream, B};
  BodyT launch_body(body);  // old-style reducers in the body take their block slots here
  li.setup = false;
  hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(B), 0, stream, seg, launch_body, f_params,
                     len);
  resources::detail::check_hip(hipGetLastError(), "forall launch");
  if constexpr (has_params)
  {
    expt::ParamMultiplexer::parampack_resolve(pol, f_params, info);  // waits for the stream
  }
  else if (!Async)
  {
    res.wait();
  }
}

template<int B, bool Async, int G, typename Iterable, typename Body, typename ForallParam>
std::enable_if_t<expt::type_traits::is_ForallParamPack<ForallParam>::value,
                 resources::EventProxy<resources::Cdna4>>
forall_impl(resources::Cdna4 res, const cdna4_exec_t<B, Async, G>& pol, Iterable&& it, Body&& body,
            ForallParam f_params)
{
  using Seg = std::decay_t<Iterable>;
  using BodyT = std::decay_t<Body>;
  namespace dd = RAJA::cdna4::detail;
  constexpr bool has_params = !expt::type_traits::is_ForallParamPack_empty<ForallParam>::value;
  const Index_type len = (Index_type)it.size();
  Seg seg(it);
  if constexpr (has_params)
  {
    launch_forall<dd::kReduceItems, dd::kReduceThreadsPerCU, B, Async, G, Seg, BodyT>(
        res, pol, seg, body, f_params, len);
  }
  else
  {
    // Does the body carry old-style reducers?  A probe copy counts them.
    dd::LaunchInfo& li = dd::launch_info();
    li = dd::LaunchInfo{};
    li.probe = true;
    {
      BodyT probe(body);
      (void)probe;
    }
    const bool reducing = li.probed > 0;
    li = dd::LaunchInfo{};
    if (reducing)
    {
      launch_forall<dd::kReduceItems, dd::kReduceThreadsPerCU, B, Async, G, Seg, BodyT>(
          res, pol, seg, body, f_params, len);
    }
    else
    {
      launch_forall<dd::kForallItems, dd::kForallThreadsPerCU, B, Async, G, Seg, BodyT>(
          res, pol, seg, body, f_params, len);
    }
  }
  return resources::EventProxy<resources::Cdna4>(res);
}

}  // namespace cdna4
}  // namespace policy

}  // namespace RAJA
