// RAJA::launch: hierarchical kernels written as teams of threads
// (include/RAJA/pattern/launch/launch_core.hpp:45-675 in the reference).
//
//   RAJA::launch<LaunchPolicy<HostPol, DevicePol>>(
//       [place | resource,] LaunchParams(Teams(...), Threads(...), shmem),
//       [expt::Reduce<op>(&target) ..., Name(...),]
//       [=] RAJA_HOST_DEVICE (RAJA::LaunchContext ctx, ValOp&... reducers) {
//         RAJA::loop<LoopPolicy<HostLoop, DeviceLoop>>(ctx, segment, [&](int i) {...});
//       });
//
// The launch policy list holds a host and a device policy; a resource (Host
// or Cdna4) or an ExecPlace picks one, and the policy list alone picks the
// host entry, as in the reference.  loop/loop_icount/tile/tile_tcount pick
// the device entry of their LoopPolicy in device code and the host entry on
// the host.  Reductions given as expt::Reduce / ReduceLoc parameters reach
// the body as ValOp& after the context and are written to their targets
// when launch returns.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <utility>

#include "RAJA/config.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/policy/host.hpp"
#include "RAJA/util/plugins.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{

enum struct ExecPlace : int
{
  HOST,
  DEVICE,
  NUM_PLACES
};

struct Teams
{
  int value[3];
  RAJA_HOST_DEVICE constexpr Teams() : value{1, 1, 1} {}
  RAJA_HOST_DEVICE constexpr Teams(int i) : value{i, 1, 1} {}
  RAJA_HOST_DEVICE constexpr Teams(int i, int j) : value{i, j, 1} {}
  RAJA_HOST_DEVICE constexpr Teams(int i, int j, int k) : value{i, j, k} {}
};

struct Threads
{
  int value[3];
  RAJA_HOST_DEVICE constexpr Threads() : value{1, 1, 1} {}
  RAJA_HOST_DEVICE constexpr Threads(int i) : value{i, 1, 1} {}
  RAJA_HOST_DEVICE constexpr Threads(int i, int j) : value{i, j, 1} {}
  RAJA_HOST_DEVICE constexpr Threads(int i, int j, int k) : value{i, j, k} {}
};

struct Lanes
{
  int value;
  RAJA_HOST_DEVICE constexpr Lanes() : value(0) {}
  RAJA_HOST_DEVICE constexpr Lanes(int i) : value(i) {}
};

struct LaunchParams
{
  Teams teams;
  Threads threads;
  size_t shared_mem_size = 0;
  LaunchParams() = default;
  LaunchParams(Teams t, Threads th, size_t shmem = 0) : teams(t), threads(th), shared_mem_size(shmem) {}
};

// Per-team context: a bump allocator over the team's dynamic shared memory
// and the team barrier.
class LaunchContext
{
public:
  size_t shared_mem_offset = 0;
  void* shared_mem_ptr = nullptr;

  RAJA_HOST_DEVICE LaunchContext() {}

  template<typename T>
  RAJA_HOST_DEVICE T* getSharedMemory(size_t count)
  {
    void* p = static_cast<char*>(shared_mem_ptr) + shared_mem_offset;
    shared_mem_offset += count * sizeof(T);
    return static_cast<T*>(p);
  }

  RAJA_HOST_DEVICE void releaseSharedMemory() { shared_mem_offset = 0; }

  RAJA_HOST_DEVICE void teamSync() const
  {
#if defined(__HIP_DEVICE_COMPILE__)
    __syncthreads();
#endif
  }
};

template<typename HOST_POLICY, typename DEVICE_POLICY = HOST_POLICY>
struct LoopPolicy
{
  using host_policy_t = HOST_POLICY;
  using device_policy_t = DEVICE_POLICY;
};

template<typename HOST_POLICY, typename DEVICE_POLICY = HOST_POLICY>
struct LaunchPolicy
{
  using host_policy_t = HOST_POLICY;
  using device_policy_t = DEVICE_POLICY;
};

namespace policy
{
namespace sequential
{
struct seq_launch_t
{
  static constexpr Policy policy = Policy::sequential;
  using resource_type = resources::Host;
};
}  // namespace sequential
}  // namespace policy
using policy::sequential::seq_launch_t;

template<typename LAUNCH_POLICY>
struct LaunchExecute;

template<typename POLICY, typename SEGMENT>
struct LoopExecute;

template<typename POLICY, typename SEGMENT>
struct LoopICountExecute;

template<typename POLICY, typename SEGMENT>
struct TileExecute;

template<typename POLICY, typename SEGMENT>
struct TileTCountExecute;

namespace detail
{
template<typename Pol, typename Res>
struct policy_for_resource;
template<typename Pol>
struct policy_for_resource<Pol, resources::Host>
{
  using type = typename Pol::host_policy_t;
};
#if defined(RAJA_ENABLE_CDNA4)
template<typename Pol>
struct policy_for_resource<Pol, resources::Cdna4>
{
  using type = typename Pol::device_policy_t;
};
#endif
}  // namespace detail

// ---- sequential (host) launch ------------------------------------------------
// include/RAJA/policy/sequential/launch.hpp in the reference: the body runs
// once with a host buffer as its shared memory.

template<>
struct LaunchExecute<seq_launch_t>
{
  template<typename Body, typename ForallParam>
  static resources::EventProxy<resources::Host> exec(resources::Host res, const LaunchParams& p,
                                                     Body&& body, ForallParam& f_params)
  {
    std::unique_ptr<char[]> shmem(new char[p.shared_mem_size > 0 ? p.shared_mem_size : 1]);
    LaunchContext ctx;
    ctx.shared_mem_ptr = shmem.get();
    const seq_launch_t pol{};
    expt::ParamMultiplexer::parampack_init(pol, f_params);
    {
      std::decay_t<Body> b(body);
      expt::invoke_body(f_params, b, ctx);
    }
    expt::ParamMultiplexer::parampack_resolve(pol, f_params);
    return resources::EventProxy<resources::Host>(res);
  }
};

template<typename SEGMENT>
struct LoopExecute<seq_exec, SEGMENT>
{
  template<typename BODY>
  static RAJA_HOST_DEVICE RAJA_INLINE void exec(LaunchContext const&, SEGMENT const& seg, BODY const& body)
  {
    const auto len = seg.size();
    for (decltype(seg.size()) i = 0; i < len; ++i) body(seg.at(i));
  }
};

template<typename SEGMENT>
struct LoopICountExecute<seq_exec, SEGMENT>
{
  template<typename BODY>
  static RAJA_HOST_DEVICE RAJA_INLINE void exec(LaunchContext const&, SEGMENT const& seg, BODY const& body)
  {
    const auto len = seg.size();
    for (decltype(seg.size()) i = 0; i < len; ++i) body(seg.at(i), i);
  }
};

template<typename SEGMENT>
struct TileExecute<seq_exec, SEGMENT>
{
  template<typename TILE_T, typename BODY>
  static RAJA_HOST_DEVICE RAJA_INLINE void exec(LaunchContext const&, TILE_T tile, SEGMENT const& seg,
                                                BODY const& body)
  {
    const auto len = seg.size();
    for (decltype(seg.size()) t = 0; t < len; t += tile) body(seg.slice(t, tile));
  }
};

template<typename SEGMENT>
struct TileTCountExecute<seq_exec, SEGMENT>
{
  template<typename TILE_T, typename BODY>
  static RAJA_HOST_DEVICE RAJA_INLINE void exec(LaunchContext const&, TILE_T tile, SEGMENT const& seg,
                                                BODY const& body)
  {
    const auto len = seg.size();
    for (decltype(seg.size()) t = 0, k = 0; t < len; t += tile, ++k) body(seg.slice(t, tile), k);
  }
};

// ---- launch entry points -----------------------------------------------------

// Launch on an explicit resource: Host runs the host policy, Cdna4 the
// device policy.
template<typename POLICY_LIST, typename Res, typename... Args,
         std::enable_if_t<type_traits::is_resource<std::decay_t<Res>>::value>* = nullptr>
resources::EventProxy<std::decay_t<Res>> launch(Res res, const LaunchParams& params, Args&&... args)
{
  static_assert(sizeof...(Args) >= 1, "launch needs a body");
  using Pol = typename detail::policy_for_resource<POLICY_LIST, std::decay_t<Res>>::type;
  auto f_params = expt::make_forall_param_pack(std::forward<Args>(args)...);
  auto&& body = expt::get_lambda(std::forward<Args>(args)...);
  const char* name = expt::get_kernel_name(f_params);
  util::PluginScope tools(util::make_context<Pol>(name ? name : ""));
  return LaunchExecute<Pol>::exec(res, params, body, f_params);
}

// Launch where a run-time place picks the policy.
template<typename POLICY_LIST, typename... Args>
void launch(ExecPlace place, const LaunchParams& params, Args&&... args)
{
  using HP = typename POLICY_LIST::host_policy_t;
  using DP = typename POLICY_LIST::device_policy_t;
  if (place == ExecPlace::HOST)
  {
    if constexpr (std::is_same<typename HP::resource_type, resources::Host>::value)
    {
      launch<POLICY_LIST>(resources::Host::get_default(), params, std::forward<Args>(args)...);
      return;
    }
  }
#if defined(RAJA_ENABLE_CDNA4)
  if (place == ExecPlace::DEVICE)
  {
    if constexpr (std::is_same<typename DP::resource_type, resources::Cdna4>::value)
    {
      launch<POLICY_LIST>(resources::Cdna4::get_default(), params, std::forward<Args>(args)...);
      return;
    }
  }
#endif
  (void)sizeof(DP);
  throw std::runtime_error("RAJA::launch: no policy for this place (or the device is not enabled)");
}

// Launch with the host entry of the policy list and its default resource.
template<typename POLICY_LIST, typename... Args>
auto launch(const LaunchParams& params, Args&&... args)
{
  using Pol = typename POLICY_LIST::host_policy_t;
  using Res = typename Pol::resource_type;
  auto f_params = expt::make_forall_param_pack(std::forward<Args>(args)...);
  auto&& body = expt::get_lambda(std::forward<Args>(args)...);
  const char* name = expt::get_kernel_name(f_params);
  util::PluginScope tools(util::make_context<Pol>(name ? name : ""));
  return LaunchExecute<Pol>::exec(Res::get_default(), params, body, f_params);
}

// ---- loops inside a launch -----------------------------------------------------

template<typename POLICY_LIST>
#if defined(__HIP_DEVICE_COMPILE__)
using loop_policy = typename POLICY_LIST::device_policy_t;
#else
using loop_policy = typename POLICY_LIST::host_policy_t;
#endif

template<typename POLICY_LIST, typename CONTEXT, typename SEGMENT, typename BODY>
RAJA_HOST_DEVICE RAJA_INLINE void loop(CONTEXT const& ctx, SEGMENT const& segment, BODY const& body)
{
  LoopExecute<loop_policy<POLICY_LIST>, SEGMENT>::exec(ctx, segment, body);
}

template<typename POLICY_LIST, typename CONTEXT, typename SEGMENT, typename BODY>
RAJA_HOST_DEVICE RAJA_INLINE void loop_icount(CONTEXT const& ctx, SEGMENT const& segment, BODY const& body)
{
  LoopICountExecute<loop_policy<POLICY_LIST>, SEGMENT>::exec(ctx, segment, body);
}

template<typename POLICY_LIST, typename CONTEXT, typename TILE_T, typename SEGMENT, typename BODY>
RAJA_HOST_DEVICE RAJA_INLINE void tile(CONTEXT const& ctx, TILE_T tile_size, SEGMENT const& segment,
                                       BODY const& body)
{
  TileExecute<loop_policy<POLICY_LIST>, SEGMENT>::exec(ctx, tile_size, segment, body);
}

template<typename POLICY_LIST, typename CONTEXT, typename TILE_T, typename SEGMENT, typename BODY>
RAJA_HOST_DEVICE RAJA_INLINE void tile_tcount(CONTEXT const& ctx, TILE_T tile_size, SEGMENT const& segment,
                                              BODY const& body)
{
  TileTCountExecute<loop_policy<POLICY_LIST>, SEGMENT>::exec(ctx, tile_size, segment, body);
}

}  // namespace RAJA
