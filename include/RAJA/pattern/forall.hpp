// RAJA::forall front-end (include/RAJA/pattern/forall.hpp:545-627 in the
// reference): forall<ExecPolicy>([resource,] iterable, params..., body).
// The trailing arguments before the body are expt::Reduce / ReduceLoc /
// Name parameters; the body receives one ValOp& per reducer after the index.
// Dispatch is by overload on the policy type (forall_impl in each policy's
// namespace, found by argument-dependent lookup).
#pragma once

#include <type_traits>
#include <utility>

#include "RAJA/config.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{

template<typename ExecPolicy, typename Res, typename Iterable, typename... Args,
         std::enable_if_t<type_traits::is_resource<std::decay_t<Res>>::value>* = nullptr>
resources::EventProxy<std::decay_t<Res>> forall(Res r, Iterable&& it, Args&&... args)
{
  static_assert(sizeof...(Args) >= 1, "forall needs a loop body");
  auto pack = detail::make_forall_param_pack(std::forward<Args>(args)...);
  auto&& body = detail::last_arg(std::forward<Args>(args)...);
  if (it.size() <= 0)
  {
    return resources::EventProxy<std::decay_t<Res>>(r);  // empty range: no launch
  }
  return forall_impl(r, ExecPolicy{}, std::forward<Iterable>(it), body, pack);
}

template<typename ExecPolicy, typename Iterable, typename... Args,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<Iterable>>::value>* = nullptr>
auto forall(Iterable&& it, Args&&... args)
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return forall<ExecPolicy>(Res::get_default(), std::forward<Iterable>(it),
                            std::forward<Args>(args)...);
}

}  // namespace RAJA
