// seq_exec and omp_parallel_for_exec forall and reducers.  Semantics of
// include/RAJA/policy/sequential/forall.hpp:58-102 and
// include/RAJA/policy/openmp/forall.hpp:60-120 in the reference:
//   * seq forall calls the body for every index in order on a copy of the
//     body, so captured reducers fold back when the copy is destroyed;
//   * omp forall privatises the body and the parameter pack per thread and
//     splits the range with a static schedule; the thread copies of the
//     parameters are folded (param_combine) in thread order after the region,
//     which makes them deterministic (the reference combines them in a
//     critical section, in whatever order threads arrive).
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

namespace expt
{
namespace detail
{
// seq_exec / omp_parallel_for_exec (and the sequential launch) parameter
// reductions (include/RAJA/policy/sequential/params/reduce.hpp:13-47 and
// include/RAJA/policy/openmp/params/reduce.hpp in the reference).  Thread
// copies are folded with the deterministic tree operator (ValLoc ties to the
// lower index) in thread order.
template<typename P>
using is_host_param_policy =
    std::integral_constant<bool, RAJA::type_traits::is_sequential_policy<P>::value ||
                                     RAJA::type_traits::is_openmp_policy<P>::value>;

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<is_host_param_policy<EXEC_POL>::value> param_init(EXEC_POL const&,
                                                                    Reducer<OP, T, VOp>& red)
{
  red.m_valop = VOp{};
}

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<is_host_param_policy<EXEC_POL>::value> param_combine(EXEC_POL const&,
                                                                       Reducer<OP, T, VOp>& out,
                                                                       const Reducer<OP, T, VOp>& in)
{
  using R = Reducer<OP, T, VOp>;
  out.m_valop.val = reducer_tree<R>{}(out.m_valop.val, in.m_valop.val);
}

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<is_host_param_policy<EXEC_POL>::value> param_combine(EXEC_POL const&,
                                                                       Reducer<OP, T, VOp>&)
{
}

template<typename EXEC_POL, typename OP, typename T, typename VOp>
std::enable_if_t<is_host_param_policy<EXEC_POL>::value> param_resolve(EXEC_POL const&,
                                                                       Reducer<OP, T, VOp>& red)
{
  red.combineTarget(red.m_valop.val);
}
}  // namespace detail
}  // namespace expt

namespace policy
{
namespace sequential
{

template<typename Iterable, typename Body, typename ForallParam>
std::enable_if_t<expt::type_traits::is_ForallParamPack<ForallParam>::value,
                 resources::EventProxy<resources::Host>>
forall_impl(resources::Host r, const seq_exec& pol, Iterable&& it, Body&& body, ForallParam f_params)
{
  const auto n = it.size();
  expt::ParamMultiplexer::parampack_init(pol, f_params);
  {
    std::decay_t<Body> b(body);
    for (decltype(it.size()) k = 0; k < n; ++k)
    {
      expt::invoke_body(f_params, b, it.at(k));
    }
  }
  expt::ParamMultiplexer::parampack_resolve(pol, f_params);
  return resources::EventProxy<resources::Host>(r);
}

}  // namespace sequential

namespace omp
{

template<typename Iterable, typename Body, typename ForallParam>
std::enable_if_t<expt::type_traits::is_ForallParamPack<ForallParam>::value,
                 resources::EventProxy<resources::Host>>
forall_impl(resources::Host r, const omp_parallel_for_exec& pol, Iterable&& it, Body&& body,
            ForallParam f_params)
{
  const auto n = it.size();
  expt::ParamMultiplexer::parampack_init(pol, f_params);
#if defined(RAJA_ENABLE_OPENMP)
  std::vector<ForallParam> per_thread(omp_get_max_threads(), f_params);
#pragma omp parallel
  {
    std::decay_t<Body> b(body);
    ForallParam& mine = per_thread[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (decltype(it.size()) k = 0; k < n; ++k)
    {
      expt::invoke_body(mine, b, it.at(k));
    }
  }
  for (auto& t : per_thread) expt::ParamMultiplexer::parampack_combine(pol, f_params, t);
#else
  {
    std::decay_t<Body> b(body);
    for (decltype(it.size()) k = 0; k < n; ++k) expt::invoke_body(f_params, b, it.at(k));
  }
#endif
  expt::ParamMultiplexer::parampack_resolve(pol, f_params);
  return resources::EventProxy<resources::Host>(r);
}

}  // namespace omp
}  // namespace policy

}  // namespace RAJA
