// seq_exec and omp_parallel_for_exec forall and reducers.  Semantics of
// include/RAJA/policy/sequential/forall.hpp:58-102 and
// include/RAJA/policy/openmp/forall.hpp:60-120 in the reference:
//   * seq forall calls the body for every index in order on a copy of the
//     body, so captured reducers fold back when the copy is destroyed;
//   * omp forall privatises the body per thread and splits the range with a
//     static schedule; parameter reductions are folded in thread order after
//     the region, which makes them deterministic (the reference combines
//     them in a critical section, in whatever order threads arrive).
#pragma once

#include <algorithm>
#include <numeric>
#include <vector>
#if defined(RAJA_ENABLE_OPENMP)
#include <omp.h>
#endif

#include "RAJA/config.hpp"
#include "RAJA/index/RangeSegment.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/pattern/reduce.hpp"
#include "RAJA/policy/PolicyBase.hpp"

namespace RAJA
{

template<typename T, typename Op>
using SeqReduceImpl = reduce::detail::HostCombinable<T, Op, false>;
template<typename T, typename Op>
using OmpReduceImpl = reduce::detail::HostCombinable<T, Op, true>;
RAJA_CDNA4_DECLARE_REDUCERS(seq_reduce, SeqReduceImpl)
RAJA_CDNA4_DECLARE_REDUCERS(omp_reduce, OmpReduceImpl)

namespace detail
{
template<typename... Ps, typename... Vs, size_t... I>
void resolve_host_params(std::tuple<Ps...>& ps, std::tuple<Vs...>& vs, std::index_sequence<I...>)
{
  (void)std::initializer_list<int>{
      0, (expt::detail::combine_target(std::get<I>(ps), std::get<I>(vs).val), 0)...};
}

template<typename... Vs, size_t... I>
void fold_valops(std::tuple<Vs...>& into, const std::tuple<Vs...>& from, std::index_sequence<I...>)
{
  (void)std::initializer_list<int>{
      0, (std::get<I>(into).val = typename expt::detail::valop_tree<Vs>::type{}(
              std::get<I>(into).val, std::get<I>(from).val),
          0)...};
}
}  // namespace detail

namespace policy
{
namespace sequential
{

template<typename Iterable, typename Body, typename Pack>
resources::EventProxy<resources::Host> forall_impl(resources::Host r, const seq_exec&,
                                                   Iterable&& it, Body&& body, Pack& pack)
{
  const auto n = it.size();
  auto vops = RAJA::detail::make_valops(pack.reducers);
  {
    std::decay_t<Body> b(body);
    for (decltype(it.size()) k = 0; k < n; ++k)
    {
      RAJA::detail::invoke_body(b, it.at(k), vops);
    }
  }
  RAJA::detail::resolve_host_params(
      pack.reducers, vops, std::make_index_sequence<std::tuple_size<decltype(vops)>::value>{});
  return resources::EventProxy<resources::Host>(r);
}

}  // namespace sequential

namespace omp
{

template<typename Iterable, typename Body, typename Pack>
resources::EventProxy<resources::Host> forall_impl(resources::Host r, const omp_parallel_for_exec&,
                                                   Iterable&& it, Body&& body, Pack& pack)
{
  using VT = decltype(RAJA::detail::make_valops(pack.reducers));
  const auto n = it.size();
#if defined(RAJA_ENABLE_OPENMP)
  std::vector<VT> per_thread(omp_get_max_threads());
#pragma omp parallel
  {
    std::decay_t<Body> b(body);
    VT vops = RAJA::detail::make_valops(pack.reducers);
#pragma omp for schedule(static)
    for (decltype(it.size()) k = 0; k < n; ++k)
    {
      RAJA::detail::invoke_body(b, it.at(k), vops);
    }
    per_thread[omp_get_thread_num()] = vops;
  }
  VT total = RAJA::detail::make_valops(pack.reducers);
  for (auto& v : per_thread)
  {
    RAJA::detail::fold_valops(total, v, std::make_index_sequence<std::tuple_size<VT>::value>{});
  }
#else
  VT total = RAJA::detail::make_valops(pack.reducers);
  {
    std::decay_t<Body> b(body);
    for (decltype(it.size()) k = 0; k < n; ++k) RAJA::detail::invoke_body(b, it.at(k), total);
  }
#endif
  RAJA::detail::resolve_host_params(pack.reducers, total,
                                    std::make_index_sequence<std::tuple_size<VT>::value>{});
  return resources::EventProxy<resources::Host>(r);
}

}  // namespace omp
}  // namespace policy

}  // namespace RAJA
