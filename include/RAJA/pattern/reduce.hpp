// Reducer objects (old style): ReduceSum/Min/Max/MinLoc/MaxLoc/BitOr/BitAnd.
//
// Declarations and semantics follow include/RAJA/pattern/reduce.hpp:79-233
// and include/RAJA/pattern/detail/reduce.hpp:24-593 in the reference:
//   * a reducer captured by value in a loop body combines its copies' values
//     into the original; get() returns op(init, contributions);
//   * reset(v) is get() followed by re-initialisation;
//   * MinLoc/MaxLoc keep (value, index); the initial pair wins ties and the
//     default index is -1;
//   * a reducer accumulates across consecutive foralls until reset.
// Each backend specialises the storage ("impl") class; the front classes
// below add the user-facing operators.
#pragma once

#include <memory>
#include <mutex>

#include "RAJA/config.hpp"
#include "RAJA/util/Operators.hpp"

namespace RAJA
{
namespace reduce
{
namespace detail
{

template<typename T, bool = std::is_integral<T>::value>
struct DefaultLoc
{
  RAJA_HOST_DEVICE static constexpr T value() { return T(); }
};
template<typename T>
struct DefaultLoc<T, true>
{
  RAJA_HOST_DEVICE static constexpr T value() { return T(-1); }
};

// (value, index) pair; comparisons look at the value only
// (include/RAJA/pattern/detail/reduce.hpp:127-178).
template<typename T, typename IndexType, bool doing_min = true>
struct ValueLoc
{
  T val;
  IndexType loc;
  RAJA_HOST_DEVICE constexpr ValueLoc()
      : val(doing_min ? operators::limits<T>::max() : operators::limits<T>::min()),
        loc(DefaultLoc<IndexType>::value())
  {}
  RAJA_HOST_DEVICE constexpr ValueLoc(T v) : val(v), loc(DefaultLoc<IndexType>::value()) {}
  RAJA_HOST_DEVICE constexpr ValueLoc(T v, IndexType l) : val(v), loc(l) {}
  RAJA_HOST_DEVICE operator T() const { return val; }
  RAJA_HOST_DEVICE IndexType getLoc() const { return loc; }
  RAJA_HOST_DEVICE bool operator<(const ValueLoc& r) const { return val < r.val; }
  RAJA_HOST_DEVICE bool operator>(const ValueLoc& r) const { return val > r.val; }
};

// Operator functors used by the reducers (left fold semantics).
template<typename T>
struct sum_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return a + b; }
  RAJA_HOST_DEVICE static T identity() { return T(0); }
};
template<typename T>
struct min_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return (b < a) ? b : a; }
  RAJA_HOST_DEVICE static T identity() { return operators::limits<T>::max(); }
};
template<typename T>
struct max_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return (a < b) ? b : a; }
  RAJA_HOST_DEVICE static T identity() { return operators::limits<T>::min(); }
};
template<typename T>
struct or_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return a | b; }
  RAJA_HOST_DEVICE static T identity() { return T(0); }
};
template<typename T>
struct and_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return a & b; }
  RAJA_HOST_DEVICE static T identity() { return T(~T(0)); }
};
template<typename T, typename I>
struct minloc_op
{
  using V = ValueLoc<T, I, true>;
  RAJA_HOST_DEVICE V operator()(const V& a, const V& b) const { return (b.val < a.val) ? b : a; }
  RAJA_HOST_DEVICE static V identity() { return V(); }
};
template<typename T, typename I>
struct maxloc_op
{
  using V = ValueLoc<T, I, false>;
  RAJA_HOST_DEVICE V operator()(const V& a, const V& b) const { return (a.val < b.val) ? b : a; }
  RAJA_HOST_DEVICE static V identity() { return V(); }
};

template<typename T>
RAJA_HOST_DEVICE inline bool differs(const T& a, const T& b)
{
  return a != b;
}
template<typename T, typename I, bool M>
RAJA_HOST_DEVICE inline bool differs(const ValueLoc<T, I, M>& a, const ValueLoc<T, I, M>& b)
{
  return a.val != b.val;
}

// Parallel fold operator of a reducer operator: (value, index) ties go to
// the lower index, which is the first one for increasing iteration order.
template<typename Op>
struct tree_of
{
  using type = Op;
};
template<typename T, typename I>
struct tree_of<minloc_op<T, I>>
{
  struct type
  {
    using V = ValueLoc<T, I, true>;
    RAJA_HOST_DEVICE V operator()(const V& a, const V& b) const
    {
      return (b.val < a.val || (b.val == a.val && b.loc < a.loc)) ? b : a;
    }
  };
};
template<typename T, typename I>
struct tree_of<maxloc_op<T, I>>
{
  struct type
  {
    using V = ValueLoc<T, I, false>;
    RAJA_HOST_DEVICE V operator()(const V& a, const V& b) const
    {
      return (a.val < b.val || (b.val == a.val && b.loc < a.loc)) ? b : a;
    }
  };
};

// Host combinable for seq_reduce and omp_reduce: copies start at the
// identity and fold into the original when destroyed, as BaseCombinable
// does (detail/reduce.hpp:282-358).  With PARALLEL (omp_reduce) the copies of
// concurrent threads fold among themselves with the tie-breaking tree
// operator under a lock, and the initial value is applied at get() (it wins
// ties), so MinLoc/MaxLoc do not depend on which thread finishes first.
template<typename T, typename Op, bool PARALLEL = false>
class HostCombinable
{
  using Tree = typename tree_of<Op>::type;

public:
  using value_type = T;
  HostCombinable(T init, T identity)
      : m_root(this), m_identity(identity), m_value(init), m_children(identity)
  {}
  HostCombinable(const HostCombinable& o)
      : m_root(o.m_root), m_identity(o.m_identity), m_value(o.m_identity),
        m_children(o.m_identity)
  {}
  ~HostCombinable()
  {
    if (m_root != this && differs(m_value, m_identity))
    {
      std::lock_guard<std::mutex> g(m_root->m_lock);
      if (PARALLEL)
      {
        m_root->m_children = Tree{}(m_root->m_children, m_value);
      }
      else
      {
        m_root->m_value = Op{}(m_root->m_value, m_value);
      }
    }
  }
  HostCombinable& operator=(const HostCombinable&) = delete;
  void combine(const T& v) const { m_value = Op{}(m_value, v); }
  T get() const
  {
    std::lock_guard<std::mutex> g(m_lock);
    if (PARALLEL && differs(m_children, m_identity))
    {
      m_value = Op{}(m_value, m_children);
      m_children = m_identity;
    }
    return m_value;
  }
  void reset(T init, T identity)
  {
    m_value = init;
    m_identity = identity;
    m_children = identity;
  }

private:
  HostCombinable* m_root;
  T m_identity;
  mutable T m_value;
  mutable T m_children;
  mutable std::mutex m_lock;
};

}  // namespace detail
}  // namespace reduce

// Primary templates; backends specialise them on their reduce policy.
template<typename REDUCE_POLICY, typename T>
class ReduceSum;
template<typename REDUCE_POLICY, typename T>
class ReduceMin;
template<typename REDUCE_POLICY, typename T>
class ReduceMax;
template<typename REDUCE_POLICY, typename T, typename IndexType = Index_type>
class ReduceMinLoc;
template<typename REDUCE_POLICY, typename T, typename IndexType = Index_type>
class ReduceMaxLoc;
template<typename REDUCE_POLICY, typename T>
class ReduceBitOr;
template<typename REDUCE_POLICY, typename T>
class ReduceBitAnd;

namespace reduce
{
namespace detail
{

// Front classes: the user-facing operators over a backend storage `Impl`
// (which provides combine(), get() and reset()).
template<typename Impl, typename T>
class SumFront : public Impl
{
public:
  SumFront(T init = T(), T identity = T(0)) : Impl(init, identity) {}
  RAJA_HOST_DEVICE const SumFront& operator+=(T v) const
  {
    this->combine(v);
    return *this;
  }
  T get() const { return Impl::get(); }
  operator T() const { return get(); }
  void reset(T init, T identity = T(0))
  {
    (void)get();
    Impl::reset(init, identity);
  }
};

template<typename Impl, typename T, typename Op>
class MinMaxFront : public Impl
{
public:
  MinMaxFront(T init = Op::identity(), T identity = Op::identity()) : Impl(init, identity) {}
  RAJA_HOST_DEVICE const MinMaxFront& min(T v) const
  {
    this->combine(v);
    return *this;
  }
  RAJA_HOST_DEVICE const MinMaxFront& max(T v) const
  {
    this->combine(v);
    return *this;
  }
  T get() const { return Impl::get(); }
  operator T() const { return get(); }
  void reset(T init, T identity = Op::identity())
  {
    (void)get();
    Impl::reset(init, identity);
  }
};

template<typename Impl, typename T, typename Op>
class BitFront : public Impl
{
public:
  BitFront(T init = Op::identity(), T identity = Op::identity()) : Impl(init, identity) {}
  RAJA_HOST_DEVICE const BitFront& operator|=(T v) const
  {
    this->combine(v);
    return *this;
  }
  RAJA_HOST_DEVICE const BitFront& operator&=(T v) const
  {
    this->combine(v);
    return *this;
  }
  T get() const { return Impl::get(); }
  operator T() const { return get(); }
  void reset(T init, T identity = Op::identity())
  {
    (void)get();
    Impl::reset(init, identity);
  }
};

template<typename Impl, typename T, typename IndexType, bool doing_min>
class LocFront : public Impl
{
  using VL = ValueLoc<T, IndexType, doing_min>;

public:
  LocFront() : Impl(VL(), VL()) {}
  LocFront(T init_val, IndexType init_loc,
           T identity_val = doing_min ? operators::limits<T>::max() : operators::limits<T>::min(),
           IndexType identity_loc = DefaultLoc<IndexType>::value())
      : Impl(VL(init_val, init_loc), VL(identity_val, identity_loc))
  {}
  RAJA_HOST_DEVICE const LocFront& minloc(T v, IndexType i) const
  {
    this->combine(VL(v, i));
    return *this;
  }
  RAJA_HOST_DEVICE const LocFront& maxloc(T v, IndexType i) const
  {
    this->combine(VL(v, i));
    return *this;
  }
  T get() const { return Impl::get().val; }
  IndexType getLoc() const { return Impl::get().loc; }
  operator T() const { return get(); }
  void reset(T init_val, IndexType init_loc,
             T identity_val = doing_min ? operators::limits<T>::max() : operators::limits<T>::min(),
             IndexType identity_loc = DefaultLoc<IndexType>::value())
  {
    (void)Impl::get();
    Impl::reset(VL(init_val, init_loc), VL(identity_val, identity_loc));
  }
};

}  // namespace detail
}  // namespace reduce

#define RAJA_CDNA4_DECLARE_REDUCERS(POLICY, IMPL)                                              \
  template<typename T>                                                                         \
  class ReduceSum<POLICY, T>                                                                   \
      : public reduce::detail::SumFront<IMPL<T, reduce::detail::sum_op<T>>, T>                 \
  {                                                                                            \
    using Base = reduce::detail::SumFront<IMPL<T, reduce::detail::sum_op<T>>, T>;              \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T>                                                                         \
  class ReduceMin<POLICY, T>                                                                   \
      : public reduce::detail::MinMaxFront<IMPL<T, reduce::detail::min_op<T>>, T,              \
                                           reduce::detail::min_op<T>>                          \
  {                                                                                            \
    using Base = reduce::detail::MinMaxFront<IMPL<T, reduce::detail::min_op<T>>, T,            \
                                             reduce::detail::min_op<T>>;                       \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T>                                                                         \
  class ReduceMax<POLICY, T>                                                                   \
      : public reduce::detail::MinMaxFront<IMPL<T, reduce::detail::max_op<T>>, T,              \
                                           reduce::detail::max_op<T>>                          \
  {                                                                                            \
    using Base = reduce::detail::MinMaxFront<IMPL<T, reduce::detail::max_op<T>>, T,            \
                                             reduce::detail::max_op<T>>;                       \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T>                                                                         \
  class ReduceBitOr<POLICY, T>                                                                 \
      : public reduce::detail::BitFront<IMPL<T, reduce::detail::or_op<T>>, T,                  \
                                        reduce::detail::or_op<T>>                              \
  {                                                                                            \
    using Base = reduce::detail::BitFront<IMPL<T, reduce::detail::or_op<T>>, T,                \
                                          reduce::detail::or_op<T>>;                           \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T>                                                                         \
  class ReduceBitAnd<POLICY, T>                                                                \
      : public reduce::detail::BitFront<IMPL<T, reduce::detail::and_op<T>>, T,                 \
                                        reduce::detail::and_op<T>>                             \
  {                                                                                            \
    using Base = reduce::detail::BitFront<IMPL<T, reduce::detail::and_op<T>>, T,               \
                                          reduce::detail::and_op<T>>;                          \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T, typename I>                                                             \
  class ReduceMinLoc<POLICY, T, I>                                                             \
      : public reduce::detail::LocFront<                                                       \
            IMPL<reduce::detail::ValueLoc<T, I, true>, reduce::detail::minloc_op<T, I>>, T, I, \
            true>                                                                              \
  {                                                                                            \
    using Base = reduce::detail::LocFront<                                                     \
        IMPL<reduce::detail::ValueLoc<T, I, true>, reduce::detail::minloc_op<T, I>>, T, I,     \
        true>;                                                                                 \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };                                                                                           \
  template<typename T, typename I>                                                             \
  class ReduceMaxLoc<POLICY, T, I>                                                             \
      : public reduce::detail::LocFront<                                                       \
            IMPL<reduce::detail::ValueLoc<T, I, false>, reduce::detail::maxloc_op<T, I>>, T,   \
            I, false>                                                                          \
  {                                                                                            \
    using Base = reduce::detail::LocFront<                                                     \
        IMPL<reduce::detail::ValueLoc<T, I, false>, reduce::detail::maxloc_op<T, I>>, T, I,    \
        false>;                                                                                \
  public:                                                                                      \
    using Base::Base;                                                                          \
  };

}  // namespace RAJA
