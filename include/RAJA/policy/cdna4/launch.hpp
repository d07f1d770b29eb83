// RAJA::launch on the cdna4 policies (include/RAJA/policy/hip/launch.hpp:30-1986
// and the hip loop policies of include/RAJA/policy/hip/policy.hpp:1600-1850
// in the reference).
//
// cdna4_launch_t<Async, NThreads> runs the body as one kernel: Teams are the
// grid, Threads the block, LaunchParams::shared_mem_size the dynamic LDS of
// each team.  NThreads > 0 fixes the kernel's launch bound (and must be the
// team size); 0 allows up to 1024 threads.  Reducer parameters fold per
// team with the same block fold as forall (blocks numbered x fastest), one
// value per team goes to pinned host slots, and the host folds them in team
// order at the end of launch -- a deterministic result for a given launch
// shape.  Old-style reducers captured by the body work as in forall.
//
// Loop policies map a loop onto one hardware index:
//   cdna4_thread_{x,y,z}_{direct,loop}   threadIdx, stride blockDim
//   cdna4_block_{x,y,z}_{direct,loop}    blockIdx, stride gridDim
//   cdna4_global_{x,y,z}_{direct,loop}   blockIdx*blockDim+threadIdx, stride grid*block
// "direct" runs iteration idx when idx < length; "loop" strides over the
// whole segment.  cdna4_global_thread_{x,y,z} name the global direct ones.
#pragma once

#include "RAJA/pattern/launch.hpp"
#include "RAJA/policy/cdna4/forall.hpp"

namespace RAJA
{
namespace policy
{
namespace cdna4
{

template<bool Async = false, int NThreads = 0>
struct cdna4_launch_t
{
  static constexpr Policy policy = Policy::cdna4;
  using resource_type = resources::Cdna4;
  static constexpr bool async = Async;
  static constexpr int num_threads = NThreads;
};

enum class hw_index
{
  thread,
  block,
  global
};

template<int Dim, hw_index Kind, bool Loop>
struct cdna4_indexer
{
  static constexpr Policy policy = Policy::cdna4;
  RAJA_DEVICE static inline Index_type index()
  {
    const unsigned t = Dim == 0 ? threadIdx.x : (Dim == 1 ? threadIdx.y : threadIdx.z);
    const unsigned b = Dim == 0 ? blockIdx.x : (Dim == 1 ? blockIdx.y : blockIdx.z);
    const unsigned bd = Dim == 0 ? blockDim.x : (Dim == 1 ? blockDim.y : blockDim.z);
    if (Kind == hw_index::thread) return (Index_type)t;
    if (Kind == hw_index::block) return (Index_type)b;
    return (Index_type)b * bd + t;
  }
  RAJA_DEVICE static inline Index_type size()
  {
    const unsigned bd = Dim == 0 ? blockDim.x : (Dim == 1 ? blockDim.y : blockDim.z);
    const unsigned gd = Dim == 0 ? gridDim.x : (Dim == 1 ? gridDim.y : gridDim.z);
    if (Kind == hw_index::thread) return (Index_type)bd;
    if (Kind == hw_index::block) return (Index_type)gd;
    return (Index_type)gd * bd;
  }
};

}  // namespace cdna4
}  // namespace policy

using policy::cdna4::cdna4_launch_t;

using cdna4_thread_x_direct = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::thread, false>;
using cdna4_thread_y_direct = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::thread, false>;
using cdna4_thread_z_direct = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::thread, false>;
using cdna4_thread_x_loop = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::thread, true>;
using cdna4_thread_y_loop = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::thread, true>;
using cdna4_thread_z_loop = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::thread, true>;
using cdna4_block_x_direct = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::block, false>;
using cdna4_block_y_direct = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::block, false>;
using cdna4_block_z_direct = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::block, false>;
using cdna4_block_x_loop = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::block, true>;
using cdna4_block_y_loop = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::block, true>;
using cdna4_block_z_loop = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::block, true>;
using cdna4_global_x_direct = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::global, false>;
using cdna4_global_y_direct = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::global, false>;
using cdna4_global_z_direct = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::global, false>;
using cdna4_global_x_loop = policy::cdna4::cdna4_indexer<0, policy::cdna4::hw_index::global, true>;
using cdna4_global_y_loop = policy::cdna4::cdna4_indexer<1, policy::cdna4::hw_index::global, true>;
using cdna4_global_z_loop = policy::cdna4::cdna4_indexer<2, policy::cdna4::hw_index::global, true>;
using cdna4_global_thread_x = cdna4_global_x_direct;
using cdna4_global_thread_y = cdna4_global_y_direct;
using cdna4_global_thread_z = cdna4_global_z_direct;

// ---- loops ------------------------------------------------------------------------
// Device-only: in host code these policies are never selected (loop picks
// the host entry of a LoopPolicy there).

template<int D, policy::cdna4::hw_index K, bool L, typename SEGMENT>
struct LoopExecute<policy::cdna4::cdna4_indexer<D, K, L>, SEGMENT>
{
  using Ix = policy::cdna4::cdna4_indexer<D, K, L>;
  template<typename BODY>
  static RAJA_DEVICE RAJA_INLINE void exec(LaunchContext const&, SEGMENT const& seg, BODY const& body)
  {
    const Index_type len = (Index_type)seg.size();
    const Index_type i0 = Ix::index();
    if (!L)
    {
      if (i0 < len) body(seg.at(i0));
      return;
    }
    const Index_type st = Ix::size();
    for (Index_type i = i0; i < len; i += st) body(seg.at(i));
  }
};

template<int D, policy::cdna4::hw_index K, bool L, typename SEGMENT>
struct LoopICountExecute<policy::cdna4::cdna4_indexer<D, K, L>, SEGMENT>
{
  using Ix = policy::cdna4::cdna4_indexer<D, K, L>;
  template<typename BODY>
  static RAJA_DEVICE RAJA_INLINE void exec(LaunchContext const&, SEGMENT const& seg, BODY const& body)
  {
    const Index_type len = (Index_type)seg.size();
    const Index_type i0 = Ix::index();
    if (!L)
    {
      if (i0 < len) body(seg.at(i0), i0);
      return;
    }
    const Index_type st = Ix::size();
    for (Index_type i = i0; i < len; i += st) body(seg.at(i), i);
  }
};

template<int D, policy::cdna4::hw_index K, bool L, typename SEGMENT>
struct TileExecute<policy::cdna4::cdna4_indexer<D, K, L>, SEGMENT>
{
  using Ix = policy::cdna4::cdna4_indexer<D, K, L>;
  template<typename TILE_T, typename BODY>
  static RAJA_DEVICE RAJA_INLINE void exec(LaunchContext const&, TILE_T tile, SEGMENT const& seg,
                                           BODY const& body)
  {
    const Index_type len = (Index_type)seg.size();
    const Index_type t0 = Ix::index() * (Index_type)tile;
    if (!L)
    {
      if (t0 < len) body(seg.slice(t0, tile));
      return;
    }
    const Index_type st = Ix::size() * (Index_type)tile;
    for (Index_type t = t0; t < len; t += st) body(seg.slice(t, tile));
  }
};

template<int D, policy::cdna4::hw_index K, bool L, typename SEGMENT>
struct TileTCountExecute<policy::cdna4::cdna4_indexer<D, K, L>, SEGMENT>
{
  using Ix = policy::cdna4::cdna4_indexer<D, K, L>;
  template<typename TILE_T, typename BODY>
  static RAJA_DEVICE RAJA_INLINE void exec(LaunchContext const&, TILE_T tile, SEGMENT const& seg,
                                           BODY const& body)
  {
    const Index_type len = (Index_type)seg.size();
    const Index_type k0 = Ix::index();
    if (!L)
    {
      if (k0 * (Index_type)tile < len) body(seg.slice(k0 * (Index_type)tile, tile), k0);
      return;
    }
    for (Index_type k = k0; k * (Index_type)tile < len; k += Ix::size())
    {
      body(seg.slice(k * (Index_type)tile, tile), k);
    }
  }
};

// ---- the launch kernel ----------------------------------------------------------------

namespace cdna4
{
namespace detail
{

template<int NT>
struct launch_bound
{
  static constexpr int value = NT > 0 ? NT : 1024;
};

template<typename Pol, int NT, typename Body, typename ForallParam>
__global__ __launch_bounds__(launch_bound<NT>::value) void launch_kernel(Body body, ForallParam f_params)
{
  extern __shared__ __attribute__((aligned(16))) char raja_cdna4_launch_shmem[];
  LaunchContext ctx;
  ctx.shared_mem_ptr = raja_cdna4_launch_shmem;
  Body b(body);  // device-private copy: its old-style reducers fold per team at scope end
  RAJA::expt::invoke_body(f_params, b, ctx);
  if constexpr (!RAJA::expt::type_traits::is_ForallParamPack_empty<ForallParam>::value)
  {
    RAJA::expt::ParamMultiplexer::parampack_combine(Pol{}, f_params);
  }
}

}  // namespace detail
}  // namespace cdna4

template<bool Async, int NT>
struct LaunchExecute<policy::cdna4::cdna4_launch_t<Async, NT>>
{
  template<typename Body, typename ForallParam>
  static resources::EventProxy<resources::Cdna4> exec(resources::Cdna4 res, const LaunchParams& p,
                                                      Body&& body, ForallParam& f_params)
  {
    using Pol = policy::cdna4::cdna4_launch_t<Async, NT>;
    using BodyT = std::decay_t<Body>;
    constexpr bool has_params = !expt::type_traits::is_ForallParamPack_empty<ForallParam>::value;
    const Pol pol{};
    const dim3 grid((unsigned)p.teams.value[0], (unsigned)p.teams.value[1], (unsigned)p.teams.value[2]);
    const dim3 block((unsigned)p.threads.value[0], (unsigned)p.threads.value[1],
                     (unsigned)p.threads.value[2]);
    const long long nteams = (long long)grid.x * grid.y * grid.z;
    const long long nthreads = (long long)block.x * block.y * block.z;
    if (p.teams.value[0] <= 0 || p.teams.value[1] <= 0 || p.teams.value[2] <= 0 ||
        p.threads.value[0] <= 0 || p.threads.value[1] <= 0 || p.threads.value[2] <= 0)
    {
      // nothing to run: parameters resolve to their initial targets
      RAJA::cdna4::detail::cdna4Info none(res, dim3(0, 0, 0), dim3(0, 0, 0));
      if constexpr (has_params)
      {
        expt::ParamMultiplexer::parampack_init(pol, f_params, none);
        expt::ParamMultiplexer::parampack_resolve(pol, f_params, none);
      }
      return resources::EventProxy<resources::Cdna4>(res);
    }
    if (nthreads > RAJA::cdna4::detail::launch_bound<NT>::value || nteams > 0x7FFFFFFFLL)
    {
      throw std::runtime_error("RAJA::launch: team size above the policy's thread bound, or too many teams");
    }
    hipStream_t stream = res.get_stream();
    RAJA::cdna4::detail::cdna4Info info(res, grid, block);
    if constexpr (has_params) expt::ParamMultiplexer::parampack_init(pol, f_params, info);
    RAJA::cdna4::detail::LaunchInfo& li = RAJA::cdna4::detail::launch_info();
    li = RAJA::cdna4::detail::LaunchInfo{true, (int)nteams, stream, (int)nthreads};
    BodyT launch_body(body);  // old-style reducers in the body take their team slots here
    li.setup = false;
    hipLaunchKernelGGL((RAJA::cdna4::detail::launch_kernel<Pol, NT, BodyT, ForallParam>), grid, block,
                       p.shared_mem_size, stream, launch_body, f_params);
    resources::detail::check_hip(hipGetLastError(), "launch");
    if constexpr (has_params)
    {
      expt::ParamMultiplexer::parampack_resolve(pol, f_params, info);  // waits for the stream
    }
    else if (!Async)
    {
      res.wait();
    }
    return resources::EventProxy<resources::Cdna4>(res);
  }
};

}  // namespace RAJA
