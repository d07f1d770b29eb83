// TypedIndexSet: an ordered collection of segments of the listed types
// (include/RAJA/index/IndexSet.hpp:50-470 in the reference), iterated by
// forall with an ExecPolicy<segment iteration, segment execution> pair:
//
//   RAJA::TypedIndexSet<RAJA::RangeSegment, RAJA::TypedListSegment<Index_type>> iset;
//   iset.push_back(RAJA::RangeSegment(0, 100));
//   iset.push_back(list);                       // copied (a list copy is a view)
//   RAJA::forall<RAJA::ExecPolicy<RAJA::seq_segit, RAJA::cdna4_exec<256>>>(iset, body);
//
// seq_segit runs the segments one after another, each with the segment
// policy; on a GPU policy every segment is a launch on the same stream, so
// segment k starts after segment k-1 has finished, as in the reference
// (include/RAJA/policy/hip/forall.hpp:659-686).  omp_parallel_segit hands
// whole segments to OpenMP threads.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <variant>
#include <vector>

#include "RAJA/config.hpp"

namespace RAJA
{

template<typename SEG_ITER_POLICY_T, typename SEG_EXEC_POLICY_T = void>
struct ExecPolicy
{
  using seg_it = SEG_ITER_POLICY_T;
  using seg_exec = SEG_EXEC_POLICY_T;
};

namespace policy
{
namespace sequential
{
struct seq_segit
{
};
}  // namespace sequential
namespace omp
{
struct omp_parallel_segit
{
};
}  // namespace omp
}  // namespace policy
using policy::omp::omp_parallel_segit;
using policy::sequential::seq_segit;

template<typename... SegTypes>
class TypedIndexSet
{
  static_assert(sizeof...(SegTypes) > 0, "an index set needs at least one segment type");
  using variant_t = std::variant<SegTypes...>;

public:
  TypedIndexSet() = default;

  template<typename Seg>
  void push_back(Seg&& s)
  {
    check_type<Seg>();
    m_segs.emplace_back(std::in_place_type<std::decay_t<Seg>>, std::forward<Seg>(s));
  }
  template<typename Seg>
  void push_front(Seg&& s)
  {
    check_type<Seg>();
    m_segs.emplace(m_segs.begin(), std::in_place_type<std::decay_t<Seg>>, std::forward<Seg>(s));
  }

  size_t getNumSegments() const { return m_segs.size(); }
  static constexpr int getNumTypes() { return (int)sizeof...(SegTypes); }

  // Total number of indices over all segments.
  size_t getLength() const
  {
    size_t n = 0;
    for (const auto& s : m_segs) n += (size_t)std::visit([](const auto& g) { return (size_t)g.size(); }, s);
    return n;
  }

  // Calls f(segment) with segment k as its own type.
  template<typename F>
  void segmentCall(size_t k, F&& f) const
  {
    std::visit(std::forward<F>(f), m_segs.at(k));
  }

  template<typename Seg>
  const Seg& getSegment(size_t k) const
  {
    return std::get<Seg>(m_segs.at(k));
  }

private:
  template<typename Seg>
  static void check_type()
  {
    static_assert((std::is_same<std::decay_t<Seg>, SegTypes>::value || ...),
                  "segment type is not one of the index set's types");
  }
  std::vector<variant_t> m_segs;
};

namespace detail
{
// Policy each segment of a seq_segit loop runs with: GPU policies are made
// asynchronous (one launch per segment on the same stream) and the loop
// waits once at the end when the user's policy is synchronous.
template<typename P>
struct segit_exec
{
  using type = P;
  static constexpr bool wait_after = false;
};
}  // namespace detail

namespace type_traits
{
template<typename T>
struct is_index_set : std::false_type
{
};
template<typename... S>
struct is_index_set<TypedIndexSet<S...>> : std::true_type
{
};
}  // namespace type_traits

}  // namespace RAJA
