// Multi (binned) reducers: MultiReduceSum/Min/Max/BitOr/BitAnd
// (include/RAJA/pattern/multi_reduce.hpp and
// include/RAJA/pattern/detail/multi_reduce.hpp:64-402 in the reference).
//
//   RAJA::MultiReduceSum<POL, double> sums(num_bins, init);   // or (container)
//   RAJA::forall<EXEC>(seg, [=](Index_type i) { sums[bin[i]] += x[i]; });
//   double s3 = sums.get(3);  sums.get_all(vec);  sums.reset();
//
// Semantics: bin b's result is op(init[b], contributions to b); results
// accumulate across loops until reset; reset() first completes a get().
// Min/Max use .min(v)/.max(v), BitOr/BitAnd use |= and &=.
//
// Storage per backend ("impl"):
//   seq / omp: one array of bins per OpenMP thread, folded in thread order
//   by get() -- deterministic for a fixed thread count and schedule;
//   cdna4: see policy/cdna4/multi_reduce.hpp.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <vector>
#if defined(RAJA_ENABLE_OPENMP)
#include <omp.h>
#endif

#include "RAJA/config.hpp"
#include "RAJA/pattern/reduce.hpp"

namespace RAJA
{

template<typename POL, typename T>
struct MultiReduceSum;
template<typename POL, typename T>
struct MultiReduceMin;
template<typename POL, typename T>
struct MultiReduceMax;
template<typename POL, typename T>
struct MultiReduceBitOr;
template<typename POL, typename T>
struct MultiReduceBitAnd;

namespace policy
{
namespace sequential
{
struct seq_multi_reduce
{
};
}  // namespace sequential
namespace omp
{
struct omp_multi_reduce
{
};
}  // namespace omp
}  // namespace policy
using policy::omp::omp_multi_reduce;
using policy::sequential::seq_multi_reduce;

namespace reduce
{
namespace detail
{

// Host storage: per-thread bin arrays (one for seq), shared by all copies
// of a reducer; thread r adds into array r.
template<typename T, typename Op, bool PARALLEL>
class HostMultiImpl
{
  struct State
  {
    std::vector<T> init;
    std::vector<std::vector<T>> per_thread;
  };

public:
  using value_type = T;

  template<typename Container>
  explicit HostMultiImpl(const Container& inits) : m_state(std::make_shared<State>())
  {
    m_state->init.assign(std::begin(inits), std::end(inits));
    clear();
  }
  HostMultiImpl(const HostMultiImpl&) = default;

  void combine(size_t bin, const T& v) const
  {
    int r = 0;
#if defined(RAJA_ENABLE_OPENMP)
    if (PARALLEL) r = omp_get_thread_num();
#endif
    T& slot = m_state->per_thread[(size_t)r][bin];
    slot = Op{}(slot, v);
  }

  T get(size_t bin) const
  {
    T r = m_state->init[bin];
    for (const auto& a : m_state->per_thread) r = Op{}(r, a[bin]);
    return r;
  }

  size_t size() const { return m_state->init.size(); }

  template<typename Container>
  void reset(const Container& inits)
  {
    m_state->init.assign(std::begin(inits), std::end(inits));
    clear();
  }

private:
  void clear()
  {
    int nt = 1;
#if defined(RAJA_ENABLE_OPENMP)
    if (PARALLEL) nt = omp_get_max_threads();
#endif
    m_state->per_thread.assign((size_t)nt, std::vector<T>(m_state->init.size(), Op::identity()));
  }

  std::shared_ptr<State> m_state;
};

// User-facing front: operator[] returns a reference bound to one bin.
template<typename Impl, typename Op>
class MultiFront
{
public:
  using value_type = typename Impl::value_type;
  using T = value_type;

  explicit MultiFront(size_t num_bins, T init = Op::identity())
      : m_impl(std::vector<T>(num_bins, init))
  {}
  template<typename Container, typename = decltype(std::begin(std::declval<const Container&>()))>
  explicit MultiFront(const Container& inits) : m_impl(inits)
  {}
  RAJA_HOST_DEVICE MultiFront(const MultiFront& o) : m_impl(o.m_impl) {}
  MultiFront& operator=(const MultiFront&) = delete;

  struct reference
  {
    const MultiFront& m;
    size_t bin;
    RAJA_HOST_DEVICE const reference& operator+=(const T& v) const
    {
      m.m_impl.combine(bin, v);
      return *this;
    }
    RAJA_HOST_DEVICE const reference& min(const T& v) const
    {
      m.m_impl.combine(bin, v);
      return *this;
    }
    RAJA_HOST_DEVICE const reference& max(const T& v) const
    {
      m.m_impl.combine(bin, v);
      return *this;
    }
    RAJA_HOST_DEVICE const reference& operator|=(const T& v) const
    {
      m.m_impl.combine(bin, v);
      return *this;
    }
    RAJA_HOST_DEVICE const reference& operator&=(const T& v) const
    {
      m.m_impl.combine(bin, v);
      return *this;
    }
    T get() const { return m.get(bin); }
  };

  RAJA_HOST_DEVICE reference operator[](size_t bin) const { return reference{*this, bin}; }

  size_t size() const { return m_impl.size(); }
  T get(size_t bin) const
  {
    if (bin >= size()) throw std::out_of_range("MultiReduce::get: bin out of range");
    return m_impl.get(bin);
  }
  template<typename Container>
  void get_all(Container& out) const
  {
    if ((size_t)std::distance(std::begin(out), std::end(out)) != size())
    {
      throw std::invalid_argument("MultiReduce::get_all: container size differs from the bin count");
    }
    size_t b = 0;
    for (auto& v : out) v = m_impl.get(b++);
  }
  // reset(): every bin back to the identity, bin count kept
  void reset() { reset(size(), Op::identity()); }
  void reset(size_t num_bins, T init = Op::identity()) { reset(std::vector<T>(num_bins, init)); }
  template<typename Container, typename = decltype(std::begin(std::declval<const Container&>()))>
  void reset(const Container& inits)
  {
    for (size_t b = 0; b < size(); ++b) (void)m_impl.get(b);  // completes outstanding work
    m_impl.reset(inits);
  }

private:
  Impl m_impl;
};

}  // namespace detail
}  // namespace reduce

#define RAJA_CDNA4_DECLARE_MULTI_REDUCERS(POLICY, IMPL)                                         \
  template<typename T>                                                                         \
  struct MultiReduceSum<POLICY, T>                                                             \
      : reduce::detail::MultiFront<IMPL<T, reduce::detail::sum_op<T>>, reduce::detail::sum_op<T>> \
  {                                                                                            \
    using reduce::detail::MultiFront<IMPL<T, reduce::detail::sum_op<T>>,                       \
                                     reduce::detail::sum_op<T>>::MultiFront;                   \
  };                                                                                           \
  template<typename T>                                                                         \
  struct MultiReduceMin<POLICY, T>                                                             \
      : reduce::detail::MultiFront<IMPL<T, reduce::detail::min_op<T>>, reduce::detail::min_op<T>> \
  {                                                                                            \
    using reduce::detail::MultiFront<IMPL<T, reduce::detail::min_op<T>>,                       \
                                     reduce::detail::min_op<T>>::MultiFront;                   \
  };                                                                                           \
  template<typename T>                                                                         \
  struct MultiReduceMax<POLICY, T>                                                             \
      : reduce::detail::MultiFront<IMPL<T, reduce::detail::max_op<T>>, reduce::detail::max_op<T>> \
  {                                                                                            \
    using reduce::detail::MultiFront<IMPL<T, reduce::detail::max_op<T>>,                       \
                                     reduce::detail::max_op<T>>::MultiFront;                   \
  };                                                                                           \
  template<typename T>                                                                         \
  struct MultiReduceBitOr<POLICY, T>                                                           \
      : reduce::detail::MultiFront<IMPL<T, reduce::detail::or_op<T>>, reduce::detail::or_op<T>>   \
  {                                                                                            \
    using reduce::detail::MultiFront<IMPL<T, reduce::detail::or_op<T>>,                        \
                                     reduce::detail::or_op<T>>::MultiFront;                    \
  };                                                                                           \
  template<typename T>                                                                         \
  struct MultiReduceBitAnd<POLICY, T>                                                          \
      : reduce::detail::MultiFront<IMPL<T, reduce::detail::and_op<T>>, reduce::detail::and_op<T>> \
  {                                                                                            \
    using reduce::detail::MultiFront<IMPL<T, reduce::detail::and_op<T>>,                       \
                                     reduce::detail::and_op<T>>::MultiFront;                   \
  };

template<typename T, typename Op>
using SeqMultiImpl = reduce::detail::HostMultiImpl<T, Op, false>;
template<typename T, typename Op>
using OmpMultiImpl = reduce::detail::HostMultiImpl<T, Op, true>;
RAJA_CDNA4_DECLARE_MULTI_REDUCERS(seq_multi_reduce, SeqMultiImpl)
RAJA_CDNA4_DECLARE_MULTI_REDUCERS(omp_multi_reduce, OmpMultiImpl)

}  // namespace RAJA
