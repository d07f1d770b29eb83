// RAJA scans: inclusive_scan, exclusive_scan and their _inplace variants,
// with or without a resource argument (include/RAJA/pattern/scan.hpp:49-441
// in the reference).  Empty ranges are a no-op.  exclusive scans take an
// initial value that defaults to the operator's identity.
//
// Backends (RAJA::impl::scan::{inclusive,exclusive}[_inplace]):
//   seq_exec              include/RAJA/policy/sequential/scan.hpp:43-153
//   omp_parallel_for_exec two-level scan of include/RAJA/policy/openmp/scan.hpp:46-181
//   cdna4_exec            the single-pass decoupled-lookback kernel of
//                         libraja_cdna4.so (pointer spans of arithmetic types
//                         with plus / minimum / maximum)
#pragma once

#include <iterator>
#include <type_traits>
#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/Operators.hpp"
#include "RAJA/util/resource.hpp"

#if defined(RAJA_ENABLE_OPENMP)
#include <omp.h>
#endif

namespace RAJA
{
namespace impl
{
namespace scan
{

// ---- sequential -------------------------------------------------------------

template<typename Iter, typename BinFn>
resources::EventProxy<resources::Host> inclusive_inplace(resources::Host r, const seq_exec&,
                                                         Iter begin, Iter end, BinFn f)
{
  if (begin == end) return resources::EventProxy<resources::Host>(r);
  auto agg = *begin;
  for (Iter i = begin + 1; i != end; ++i)
  {
    agg = f(agg, *i);
    *i = agg;
  }
  return resources::EventProxy<resources::Host>(r);
}

template<typename Iter, typename BinFn, typename T>
resources::EventProxy<resources::Host> exclusive_inplace(resources::Host r, const seq_exec&,
                                                         Iter begin, Iter end, BinFn f, T v)
{
  using V = typename std::remove_reference<decltype(*begin)>::type;
  V agg = v;
  for (Iter i = begin; i != end; ++i)
  {
    V t = *i;
    *i = agg;
    agg = f(agg, t);
  }
  return resources::EventProxy<resources::Host>(r);
}

template<typename Iter, typename OutIter, typename BinFn>
resources::EventProxy<resources::Host> inclusive(resources::Host r, const seq_exec& p, Iter begin,
                                                 Iter end, OutIter out, BinFn f)
{
  std::copy(begin, end, out);
  return inclusive_inplace(r, p, out, out + (end - begin), f);
}

template<typename Iter, typename OutIter, typename BinFn, typename T>
resources::EventProxy<resources::Host> exclusive(resources::Host r, const seq_exec& p, Iter begin,
                                                 Iter end, OutIter out, BinFn f, T v)
{
  std::copy(begin, end, out);
  return exclusive_inplace(r, p, out, out + (end - begin), f, v);
}

// ---- OpenMP: per-thread chunk scans, a scan of chunk totals, a carry pass ----

template<typename Iter, typename BinFn>
resources::EventProxy<resources::Host> inclusive_inplace(resources::Host r,
                                                         const omp_parallel_for_exec&,
                                                         Iter begin, Iter end, BinFn f)
{
#if defined(RAJA_ENABLE_OPENMP)
  using V = typename std::remove_reference<decltype(*begin)>::type;
  const std::ptrdiff_t n = end - begin;
  const int p0 = (int)std::min<std::ptrdiff_t>(n, omp_get_max_threads());
  if (p0 <= 1) return inclusive_inplace(r, seq_exec{}, begin, end, f);
  std::vector<V> sums(p0, BinFn::identity());
#pragma omp parallel num_threads(p0)
  {
    const int p = omp_get_num_threads(), t = omp_get_thread_num();
    const std::ptrdiff_t b = (std::ptrdiff_t)((size_t)n * t / p);
    const std::ptrdiff_t e = (std::ptrdiff_t)((size_t)n * (t + 1) / p);
    if (b != e)
    {
      inclusive_inplace(r, seq_exec{}, begin + b, begin + e, f);
      sums[t] = begin[e - 1];
    }
#pragma omp barrier
#pragma omp single
    exclusive_inplace(r, seq_exec{}, sums.data(), sums.data() + p, f, BinFn::identity());
    for (std::ptrdiff_t i = b; i < e; ++i) begin[i] = f(sums[t], begin[i]);
  }
  return resources::EventProxy<resources::Host>(r);
#else
  return inclusive_inplace(r, seq_exec{}, begin, end, f);
#endif
}

template<typename Iter, typename BinFn, typename T>
resources::EventProxy<resources::Host> exclusive_inplace(resources::Host r,
                                                         const omp_parallel_for_exec&,
                                                         Iter begin, Iter end, BinFn f, T v)
{
#if defined(RAJA_ENABLE_OPENMP)
  using V = typename std::remove_reference<decltype(*begin)>::type;
  const std::ptrdiff_t n = end - begin;
  const int p0 = (int)std::min<std::ptrdiff_t>(n, omp_get_max_threads());
  if (p0 <= 1) return exclusive_inplace(r, seq_exec{}, begin, end, f, v);
  std::vector<V> sums(p0, BinFn::identity());
#pragma omp parallel num_threads(p0)
  {
    const int p = omp_get_num_threads(), t = omp_get_thread_num();
    const std::ptrdiff_t b = (std::ptrdiff_t)((size_t)n * t / p);
    const std::ptrdiff_t e = (std::ptrdiff_t)((size_t)n * (t + 1) / p);
    V agg = (t == 0) ? V(v) : BinFn::identity();
    for (std::ptrdiff_t i = b; i < e; ++i)
    {
      V x = begin[i];
      begin[i] = agg;
      agg = f(agg, x);
    }
    sums[t] = agg;
#pragma omp barrier
#pragma omp single
    exclusive_inplace(r, seq_exec{}, sums.data(), sums.data() + p, f, BinFn::identity());
    if (t > 0)
      for (std::ptrdiff_t i = b; i < e; ++i) begin[i] = f(sums[t], begin[i]);
  }
  return resources::EventProxy<resources::Host>(r);
#else
  return exclusive_inplace(r, seq_exec{}, begin, end, f, v);
#endif
}

template<typename Iter, typename OutIter, typename BinFn>
resources::EventProxy<resources::Host> inclusive(resources::Host r, const omp_parallel_for_exec& p,
                                                 Iter begin, Iter end, OutIter out, BinFn f)
{
  std::copy(begin, end, out);
  return inclusive_inplace(r, p, out, out + (end - begin), f);
}

template<typename Iter, typename OutIter, typename BinFn, typename T>
resources::EventProxy<resources::Host> exclusive(resources::Host r, const omp_parallel_for_exec& p,
                                                 Iter begin, Iter end, OutIter out, BinFn f, T v)
{
  std::copy(begin, end, out);
  return exclusive_inplace(r, p, out, out + (end - begin), f, v);
}

}  // namespace scan
}  // namespace impl

// ---- front-end ------------------------------------------------------------------

namespace detail
{
template<typename C>
using container_value_t = typename std::remove_cv<
    typename std::remove_reference<decltype(*std::begin(std::declval<C&>()))>::type>::type;
}  // namespace detail

template<typename ExecPolicy, typename Res, typename InContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>
resources::EventProxy<Res> inclusive_scan_inplace(Res r, InContainer&& in, Function f = Function{})
{
  auto b = std::begin(in);
  auto e = std::end(in);
  if (b == e) return resources::EventProxy<Res>(r);
  return impl::scan::inclusive_inplace(r, ExecPolicy{}, b, e, f);
}

template<typename ExecPolicy, typename InContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<InContainer>>::value>* = nullptr>
auto inclusive_scan_inplace(InContainer&& in, Function f = Function{})
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return inclusive_scan_inplace<ExecPolicy>(Res::get_default(), std::forward<InContainer>(in), f);
}

template<typename ExecPolicy, typename Res, typename InContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         typename T = detail::container_value_t<InContainer>,
         std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>
resources::EventProxy<Res> exclusive_scan_inplace(Res r, InContainer&& in, Function f = Function{},
                                                  T init = Function::identity())
{
  auto b = std::begin(in);
  auto e = std::end(in);
  if (b == e) return resources::EventProxy<Res>(r);
  return impl::scan::exclusive_inplace(r, ExecPolicy{}, b, e, f, init);
}

template<typename ExecPolicy, typename InContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         typename T = detail::container_value_t<InContainer>,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<InContainer>>::value>* = nullptr>
auto exclusive_scan_inplace(InContainer&& in, Function f = Function{}, T init = Function::identity())
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return exclusive_scan_inplace<ExecPolicy>(Res::get_default(), std::forward<InContainer>(in), f,
                                            init);
}

template<typename ExecPolicy, typename Res, typename InContainer, typename OutContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>
resources::EventProxy<Res> inclusive_scan(Res r, InContainer&& in, OutContainer&& out,
                                          Function f = Function{})
{
  auto b = std::begin(in);
  auto e = std::end(in);
  if (b == e) return resources::EventProxy<Res>(r);
  return impl::scan::inclusive(r, ExecPolicy{}, b, e, std::begin(out), f);
}

template<typename ExecPolicy, typename InContainer, typename OutContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<InContainer>>::value>* = nullptr>
auto inclusive_scan(InContainer&& in, OutContainer&& out, Function f = Function{})
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return inclusive_scan<ExecPolicy>(Res::get_default(), std::forward<InContainer>(in),
                                    std::forward<OutContainer>(out), f);
}

template<typename ExecPolicy, typename Res, typename InContainer, typename OutContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         typename T = detail::container_value_t<InContainer>,
         std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>
resources::EventProxy<Res> exclusive_scan(Res r, InContainer&& in, OutContainer&& out,
                                          Function f = Function{}, T init = Function::identity())
{
  auto b = std::begin(in);
  auto e = std::end(in);
  if (b == e) return resources::EventProxy<Res>(r);
  return impl::scan::exclusive(r, ExecPolicy{}, b, e, std::begin(out), f, init);
}

template<typename ExecPolicy, typename InContainer, typename OutContainer,
         typename Function = operators::plus<detail::container_value_t<InContainer>>,
         typename T = detail::container_value_t<InContainer>,
         std::enable_if_t<!type_traits::is_resource<std::decay_t<InContainer>>::value>* = nullptr>
auto exclusive_scan(InContainer&& in, OutContainer&& out, Function f = Function{},
                    T init = Function::identity())
{
  using Res = typename resources::get_resource<ExecPolicy>::type;
  return exclusive_scan<ExecPolicy>(Res::get_default(), std::forward<InContainer>(in),
                                    std::forward<OutContainer>(out), f, init);
}

}  // namespace RAJA
