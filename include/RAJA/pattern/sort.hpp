// RAJA sorts: sort, stable_sort, sort_pairs, stable_sort_pairs, with or
// without a resource (include/RAJA/pattern/sort.hpp:49-453 in the
// reference).  Ranges of length 0 or 1 are a no-op.  The comparator defaults
// to operators::less; operators::greater sorts descending; pairs compare
// keys only and values follow their keys.
//
// Backends (RAJA::impl::sort::{unstable,stable}[_pairs]):
//   seq_exec / omp_parallel_for_exec  std::stable_sort (a stable result is a
//                                     valid result of the unstable sorts)
//   cdna4_exec                        libraja_cdna4.so's onesweep LSD radix
//                                     sort (stable) for pointer spans of
//                                     arithmetic keys with less / greater
#pragma once

#include <algorithm>
#include <iterator>
#include <numeric>
#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/Operators.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{
namespace impl
{
namespace sort
{

template<typename Iter, typename Compare>
void host_stable(Iter begin, Iter end, Compare comp)
{
  std::stable_sort(begin, end, comp);
}

template<typename KeyIter, typename ValIter, typename Compare>
void host_stable_pairs(KeyIter kb, KeyIter ke, ValIter vb, Compare comp)
{
  const std::ptrdiff_t n = ke - kb;
  std::vector<std::ptrdiff_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](std::ptrdiff_t a, std::ptrdiff_t b) { return comp(kb[a], kb[b]); });
  using K = typename std::remove_reference<decltype(*kb)>::type;
  using V = typename std::remove_reference<decltype(*vb)>::type;
  std::vector<K> ks(n);
  std::vector<V> vs(n);
  for (std::ptrdiff_t i = 0; i < n; ++i)
  {
    ks[i] = kb[idx[i]];
    vs[i] = vb[idx[i]];
  }
  std::copy(ks.begin(), ks.end(), kb);
  std::copy(vs.begin(), vs.end(), vb);
}

#define RAJA_CDNA4_HOST_SORTS(POL)                                                            \
  template<typename Iter, typename Compare>                                                   \
  resources::EventProxy<resources::Host> unstable(resources::Host r, const POL&, Iter b,      \
                                                  Iter e, Compare c)                          \
  {                                                                                           \
    host_stable(b, e, c);                                                                     \
    return resources::EventProxy<resources::Host>(r);                                         \
  }                                                                                           \
  template<typename Iter, typename Compare>                                                   \
  resources::EventProxy<resources::Host> stable(resources::Host r, const POL&, Iter b, Iter e, \
                                                Compare c)                                    \
  {                                                                                           \
    host_stable(b, e, c);                                                                     \
    return resources::EventProxy<resources::Host>(r);                                         \
  }                                                                                           \
  template<typename KIter, typename VIter, typename Compare>                                  \
  resources::EventProxy<resources::Host> unstable_pairs(resources::Host r, const POL&,        \
                                                        KIter kb, KIter ke, VIter vb,         \
                                                        Compare c)                            \
  {                                                                                           \
    host_stable_pairs(kb, ke, vb, c);                                                         \
    return resources::EventProxy<resources::Host>(r);                                         \
  }                                                                                           \
  template<typename KIter, typename VIter, typename Compare>                                  \
  resources::EventProxy<resources::Host> stable_pairs(resources::Host r, const POL&, KIter kb, \
                                                      KIter ke, VIter vb, Compare c)          \
  {                                                                                           \
    host_stable_pairs(kb, ke, vb, c);                                                         \
    return resources::EventProxy<resources::Host>(r);                                         \
  }

RAJA_CDNA4_HOST_SORTS(seq_exec)
RAJA_CDNA4_HOST_SORTS(omp_parallel_for_exec)
#undef RAJA_CDNA4_HOST_SORTS

}  // namespace sort
}  // namespace impl

namespace detail
{
template<typename C>
using sort_value_t = typename std::remove_cv<
    typename std::remove_reference<decltype(*std::begin(std::declval<C&>()))>::type>::type;
}  // namespace detail

#define RAJA_CDNA4_SORT_FRONT(NAME, IMPL)                                                     \
  template<typename ExecPolicy, typename Res, typename Container,                             \
           typename Compare = operators::less<detail::sort_value_t<Container>>,               \
           std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>                 \
  resources::EventProxy<Res> NAME(Res r, Container&& c, Compare comp = Compare{})             \
  {                                                                                           \
    auto b = std::begin(c);                                                                   \
    auto e = std::end(c);                                                                     \
    if (e - b <= 1) return resources::EventProxy<Res>(r);                                     \
    return impl::sort::IMPL(r, ExecPolicy{}, b, e, comp);                                     \
  }                                                                                           \
  template<typename ExecPolicy, typename Container,                                           \
           typename Compare = operators::less<detail::sort_value_t<Container>>,               \
           std::enable_if_t<!type_traits::is_resource<std::decay_t<Container>>::value>* =     \
               nullptr>                                                                       \
  auto NAME(Container&& c, Compare comp = Compare{})                                          \
  {                                                                                           \
    using Res = typename resources::get_resource<ExecPolicy>::type;                           \
    return NAME<ExecPolicy>(Res::get_default(), std::forward<Container>(c), comp);            \
  }

#define RAJA_CDNA4_SORT_PAIRS_FRONT(NAME, IMPL)                                               \
  template<typename ExecPolicy, typename Res, typename KeyContainer, typename ValContainer,    \
           typename Compare = operators::less<detail::sort_value_t<KeyContainer>>,            \
           std::enable_if_t<type_traits::is_resource<Res>::value>* = nullptr>                 \
  resources::EventProxy<Res> NAME(Res r, KeyContainer&& keys, ValContainer&& vals,            \
                                  Compare comp = Compare{})                                   \
  {                                                                                           \
    auto b = std::begin(keys);                                                                \
    auto e = std::end(keys);                                                                  \
    if (e - b <= 1) return resources::EventProxy<Res>(r);                                     \
    return impl::sort::IMPL(r, ExecPolicy{}, b, e, std::begin(vals), comp);                   \
  }                                                                                           \
  template<typename ExecPolicy, typename KeyContainer, typename ValContainer,                 \
           typename Compare = operators::less<detail::sort_value_t<KeyContainer>>,            \
           std::enable_if_t<!type_traits::is_resource<std::decay_t<KeyContainer>>::value>* =  \
               nullptr>                                                                       \
  auto NAME(KeyContainer&& keys, ValContainer&& vals, Compare comp = Compare{})               \
  {                                                                                           \
    using Res = typename resources::get_resource<ExecPolicy>::type;                           \
    return NAME<ExecPolicy>(Res::get_default(), std::forward<KeyContainer>(keys),             \
                            std::forward<ValContainer>(vals), comp);                          \
  }

RAJA_CDNA4_SORT_FRONT(sort, unstable)
RAJA_CDNA4_SORT_FRONT(stable_sort, stable)
RAJA_CDNA4_SORT_PAIRS_FRONT(sort_pairs, unstable_pairs)
RAJA_CDNA4_SORT_PAIRS_FRONT(stable_sort_pairs, stable_pairs)
#undef RAJA_CDNA4_SORT_FRONT
#undef RAJA_CDNA4_SORT_PAIRS_FRONT

}  // namespace RAJA
