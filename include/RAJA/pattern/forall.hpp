// RAJA::forall front-end (include/RAJA/pattern/forall.hpp:545-627 in the
// reference): forall<ExecPolicy>([resource,] iterable, params..., body).
// The trailing arguments before the body are expt::Reduce / ReduceLoc /
// Name parameters, gathered into an expt::ForallParamPack that each
// backend's forall_impl takes by value (include/RAJA/policy/hip/forall.hpp:
// 543-560 in the reference); the body receives one ValOp& per reducer after
// the index.  Dispatch is by overload on the policy type (forall_impl in each
// policy's namespace, found by argument-dependent lookup).
#pragma once

#include <type_traits>
#include <utility>

#include "RAJA/config.hpp"
#include "RAJA/index/IndexSet.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/plugins.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{

template<typename ExecPolicy, typename Res, typename Iterable, typename... Args,
         std::enable_if_t<type_traits::is_resource<std::decay_t<Res>>::value &&
                          !type_traits::is_index_set<std::decay_t<Iterable>>::value>* = nullptr>
resources::EventProxy<std::decay_t<Res>> forall(Res r, Iterable&& it, Args&&... args)
{
  static_assert(sizeof...(Args) >= 1, "forall needs a loop body");
  auto f_params = expt::make_forall_param_pack(std::forward<Args>(args)...);
  auto&& body = expt::get_lambda(std::forward<Args>(args)...);
  const char* name = expt::get_kernel_name(f_params);
  util::PluginScope tools(util::make_context<ExecPolicy>(name ? name : ""));
  if (it.size() <= 0)
  {
    return resources::EventProxy<std::decay_t<Res>>(r);  // empty range: no launch
  }
  return forall_impl(r, ExecPolicy{}, std::forward<Iterable>(it), body, std::move(f_params));
}

template<typename ExecPolicy, typename Iterable, typename... Args,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<Iterable>>::value &&
                          !type_traits::is_index_set<std::decay_t<Iterable>>::value>* = nullptr>
auto forall(Iterable&& it, Args&&... args)
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return forall<ExecPolicy>(Res::get_default(), std::forward<Iterable>(it),
                            std::forward<Args>(args)...);
}

// ---- index sets (include/RAJA/policy/hip/forall.hpp:659-686,
// include/RAJA/policy/sequential/forall.hpp and openmp/forall.hpp in the
// reference): forall<ExecPolicy<segment iteration, segment execution>>.

namespace detail
{
template<typename SegExec, typename Res, typename... Segs, typename Body>
void forall_segments(Res r, seq_segit, const TypedIndexSet<Segs...>& iset, Body&& body)
{
  using Each = typename detail::segit_exec<SegExec>::type;
  for (size_t k = 0; k < iset.getNumSegments(); ++k)
  {
    iset.segmentCall(k, [&](const auto& seg) {
      if (seg.size() > 0) forall_impl(r, Each{}, seg, body, expt::make_forall_param_pack(body));
    });
  }
  if (detail::segit_exec<SegExec>::wait_after) r.wait();
}

template<typename SegExec, typename Res, typename... Segs, typename Body>
void forall_segments(Res r, omp_parallel_segit, const TypedIndexSet<Segs...>& iset, Body&& body)
{
  static_assert(std::is_same<Res, resources::Host>::value,
                "omp_parallel_segit runs host segment policies");
  const long ns = (long)iset.getNumSegments();
#if defined(RAJA_ENABLE_OPENMP)
#pragma omp parallel for schedule(static)
#endif
  for (long k = 0; k < ns; ++k)
  {
    iset.segmentCall((size_t)k, [&](const auto& seg) {
      if (seg.size() > 0) forall_impl(r, SegExec{}, seg, body, expt::make_forall_param_pack(body));
    });
  }
}
}  // namespace detail

template<typename ExecPolicy, typename Res, typename... Segs, typename Body,
         std::enable_if_t<type_traits::is_resource<std::decay_t<Res>>::value>* = nullptr>
resources::EventProxy<std::decay_t<Res>> forall(Res r, const TypedIndexSet<Segs...>& iset, Body&& body)
{
  util::PluginScope tools(util::make_context<ExecPolicy>(""));
  detail::forall_segments<typename ExecPolicy::seg_exec>(r, typename ExecPolicy::seg_it{}, iset,
                                                        std::forward<Body>(body));
  return resources::EventProxy<std::decay_t<Res>>(r);
}

template<typename ExecPolicy, typename... Segs, typename Body>
auto forall(const TypedIndexSet<Segs...>& iset, Body&& body)
{
  using Res = typename resources::get_resource<typename ExecPolicy::seg_exec>::type;
  return forall<ExecPolicy>(Res::get_default(), iset, std::forward<Body>(body));
}

}  // namespace RAJA
