// Parameter-based reductions: RAJA::expt::Reduce / ReduceLoc, ValLoc, ValOp,
// ValLocOp and RAJA::Name, with the semantics of
// include/RAJA/pattern/params/params_base.hpp:9-253 and
// include/RAJA/pattern/params/reducer.hpp:68-248 in the reference:
//   * the loop body receives one ValOp& per Reduce argument, in order;
//   * ValOp starts at the operator's identity; min/max update on a strict
//     comparison, so the first extremum in iteration order is kept;
//   * at the end the target becomes op(*target, result) (combineTarget).
#pragma once

#include <tuple>
#include <type_traits>
#include <utility>

#include "RAJA/config.hpp"
#include "RAJA/util/Operators.hpp"

namespace RAJA
{

// Kernel name for tools (include/RAJA/pattern/params/kernel_name.hpp).
struct Name
{
  const char* name;
  explicit Name(const char* n) : name(n) {}
};

namespace expt
{

template<typename T, typename IndexType = Index_type>
struct ValLoc
{
  using index_type = IndexType;
  using value_type = T;
  ValLoc() = default;
  RAJA_HOST_DEVICE constexpr explicit ValLoc(value_type v) : val(v) {}
  RAJA_HOST_DEVICE constexpr ValLoc(value_type v, index_type l) : val(v), loc(l) {}
  RAJA_HOST_DEVICE constexpr bool operator<(const ValLoc& rhs) const { return val < rhs.val; }
  RAJA_HOST_DEVICE constexpr bool operator>(const ValLoc& rhs) const { return val > rhs.val; }
  RAJA_HOST_DEVICE constexpr const value_type& getVal() const { return val; }
  RAJA_HOST_DEVICE constexpr const index_type& getLoc() const { return loc; }
  RAJA_HOST_DEVICE void set(T v, IndexType l)
  {
    val = v;
    loc = l;
  }
  RAJA_HOST_DEVICE void setVal(T v) { val = v; }
  RAJA_HOST_DEVICE void setLoc(IndexType l) { loc = l; }
  value_type val;
  index_type loc = -1;
};

namespace detail
{
template<typename T>
struct is_valloc : std::false_type
{
};
template<typename T, typename I>
struct is_valloc<ValLoc<T, I>> : std::true_type
{
};

// Identity of an operator for plain values and for ValLoc values.
template<template<typename, typename, typename> class Op, typename T, typename Enable = void>
struct identity_of
{
  RAJA_HOST_DEVICE static constexpr T value() { return Op<T, T, T>::identity(); }
};
template<template<typename, typename, typename> class Op, typename T, typename I>
struct identity_of<Op, ValLoc<T, I>>
{
  RAJA_HOST_DEVICE static constexpr ValLoc<T, I> value()
  {
    return ValLoc<T, I>(Op<T, T, T>::identity(), I(-1));
  }
};
}  // namespace detail

template<typename T, template<typename, typename, typename> class Op>
struct ValOp
{
  using value_type = T;
  using op_type = Op<T, T, T>;
  RAJA_HOST_DEVICE constexpr ValOp() : val(op_type::identity()) {}
  RAJA_HOST_DEVICE constexpr explicit ValOp(value_type v) : val(v) {}

  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::minimum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& min(value_type v)
  {
    if (v < val) val = v;
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::maximum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& max(value_type v)
  {
    if (v > val) val = v;
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::plus<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& operator+=(const value_type& rhs)
  {
    val += rhs;
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::bit_and<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& operator&=(const value_type& rhs)
  {
    val &= rhs;
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::bit_or<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& operator|=(const value_type& rhs)
  {
    val |= rhs;
    return *this;
  }
  value_type val;
};

template<typename T, typename IndexType, template<typename, typename, typename> class Op>
struct ValOp<ValLoc<T, IndexType>, Op>
{
  using index_type = IndexType;
  using value_type = ValLoc<T, IndexType>;
  using op_type = Op<T, T, T>;
  RAJA_HOST_DEVICE constexpr ValOp() : val(op_type::identity(), IndexType(-1)) {}
  RAJA_HOST_DEVICE constexpr explicit ValOp(value_type v) : val(v) {}
  RAJA_HOST_DEVICE constexpr ValOp(T v, IndexType l) : val(v, l) {}

  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::minimum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& minloc(T v, IndexType l)
  {
    if (v < val.val) val = value_type(v, l);
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::maximum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& maxloc(T v, IndexType l)
  {
    if (v > val.val) val = value_type(v, l);
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::minimum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& min(value_type v)
  {
    if (v < val) val = v;
    return *this;
  }
  template<typename U = op_type,
           std::enable_if_t<std::is_same<U, operators::maximum<T, T, T>>::value>* = nullptr>
  RAJA_HOST_DEVICE ValOp& max(value_type v)
  {
    if (v > val) val = v;
    return *this;
  }
  value_type val;
};

template<typename T, typename IndexType, template<typename, typename, typename> class Op>
using ValLocOp = ValOp<ValLoc<T, IndexType>, Op>;

namespace detail
{

// One Reduce(&target) argument.  value_type is what the body's ValOp holds.
template<template<typename, typename, typename> class Op, typename T>
struct Reducer
{
  using value_type = T;
  using valop_type = ValOp<T, Op>;
  T* target;
  void* device_partials = nullptr;  // backend scratch (cdna4: per-block values)
  int nblocks = 0;
};

// ReduceLoc(&val, &loc): a ValLoc reduction written back to two targets.
template<template<typename, typename, typename> class Op, typename T, typename IndexType>
struct ReducerLoc
{
  using value_type = ValLoc<T, IndexType>;
  using valop_type = ValOp<value_type, Op>;
  T* target_val;
  IndexType* target_loc;
  void* device_partials = nullptr;
  int nblocks = 0;
};

template<typename P>
struct is_reducer : std::false_type
{
};
template<template<typename, typename, typename> class Op, typename T>
struct is_reducer<Reducer<Op, T>> : std::true_type
{
};
template<template<typename, typename, typename> class Op, typename T, typename I>
struct is_reducer<ReducerLoc<Op, T, I>> : std::true_type
{
};

// Combine a finished ValOp value into the user's target (combineTarget).
template<template<typename, typename, typename> class Op, typename T>
void combine_target(Reducer<Op, T>& r, const T& v)
{
  *r.target = Op<T, T, T>{}(*r.target, v);
}
template<template<typename, typename, typename> class Op, typename T, typename I>
void combine_target(ReducerLoc<Op, T, I>& r, const ValLoc<T, I>& v)
{
  ValLoc<T, I> cur(*r.target_val, *r.target_loc);
  // operators on ValLoc compare values only and keep the left (target) on ties
  const ValLoc<T, I> res = Op<ValLoc<T, I>, ValLoc<T, I>, ValLoc<T, I>>{}(cur, v);
  *r.target_val = res.val;
  *r.target_loc = res.loc;
}

// Deterministic combine used by parallel folds: ValLoc ties go to the lower
// index (the seq result when locations increase with iteration order).
template<template<typename, typename, typename> class Op, typename T>
struct tree_op
{
  RAJA_HOST_DEVICE T operator()(const T& a, const T& b) const { return Op<T, T, T>{}(a, b); }
};
template<template<typename, typename, typename> class Op, typename T, typename I>
struct tree_op<Op, ValLoc<T, I>>
{
  RAJA_HOST_DEVICE ValLoc<T, I> operator()(const ValLoc<T, I>& a, const ValLoc<T, I>& b) const
  {
    const T r = Op<T, T, T>{}(a.val, b.val);
    if (a.val == b.val) return (b.loc < a.loc) ? b : a;
    return (r == a.val) ? a : b;
  }
};

// The parallel combine of a ValOp type.
template<typename VOp>
struct valop_tree;
template<typename T, template<typename, typename, typename> class Op>
struct valop_tree<ValOp<T, Op>>
{
  using type = tree_op<Op, T>;
};

}  // namespace detail

template<template<typename, typename, typename> class Op, typename T>
detail::Reducer<Op, T> Reduce(T* target)
{
  return detail::Reducer<Op, T>{target};
}

template<template<typename, typename, typename> class Op, typename T, typename IndexType>
detail::ReducerLoc<Op, T, IndexType> ReduceLoc(T* target_val, IndexType* target_loc)
{
  return detail::ReducerLoc<Op, T, IndexType>{target_val, target_loc};
}

}  // namespace expt

namespace detail
{

// Split forall's trailing arguments into (reducer params tuple, name, body).
template<typename... Ps>
struct ParamPack
{
  std::tuple<Ps...> reducers;
  const char* name = nullptr;
};

template<typename Pack>
RAJA_INLINE auto add_param(Pack&& pack, const Name& n)
{
  auto p = std::forward<Pack>(pack);
  p.name = n.name;
  return p;
}

template<typename... Ps, typename P,
         std::enable_if_t<expt::detail::is_reducer<std::decay_t<P>>::value>* = nullptr>
RAJA_INLINE auto add_param(ParamPack<Ps...>&& pack, P&& param)
{
  ParamPack<Ps..., std::decay_t<P>> out{
      std::tuple_cat(std::move(pack.reducers), std::make_tuple(std::forward<P>(param))),
      pack.name};
  return out;
}

template<typename Pack>
RAJA_INLINE Pack build_params(Pack&& pack)
{
  return std::forward<Pack>(pack);
}

template<typename Pack, typename First, typename... Rest>
RAJA_INLINE auto build_params(Pack&& pack, First&& first, Rest&&... rest)
{
  return build_params(add_param(std::forward<Pack>(pack), std::forward<First>(first)),
                      std::forward<Rest>(rest)...);
}

// The last argument of a parameter list.
template<typename... Args>
RAJA_INLINE decltype(auto) last_arg(Args&&... args)
{
  return std::get<sizeof...(Args) - 1>(std::forward_as_tuple(std::forward<Args>(args)...));
}

template<size_t... I, typename Tuple>
RAJA_INLINE auto make_params_from(std::index_sequence<I...>, Tuple&& t)
{
  return build_params(ParamPack<>{}, std::get<I>(std::forward<Tuple>(t))...);
}

// (params..., body) -> ParamPack of the params.
template<typename... Args>
RAJA_INLINE auto make_forall_param_pack(Args&&... args)
{
  return make_params_from(std::make_index_sequence<sizeof...(Args) - 1>{},
                          std::forward_as_tuple(std::forward<Args>(args)...));
}

// Fresh ValOps (at the identity) for every reducer of a pack.
template<typename... Ps>
RAJA_HOST_DEVICE auto make_valops(const std::tuple<Ps...>&)
{
  return std::tuple<typename Ps::valop_type...>{};
}

template<typename Body, typename Idx, typename VTuple, size_t... I>
RAJA_HOST_DEVICE RAJA_INLINE void invoke_with(Body& body, Idx i, VTuple& v, std::index_sequence<I...>)
{
  body(i, std::get<I>(v)...);
}

template<typename Body, typename Idx, typename... Vs>
RAJA_HOST_DEVICE RAJA_INLINE void invoke_body(Body& body, Idx i, std::tuple<Vs...>& v)
{
  invoke_with(body, i, v, std::index_sequence_for<Vs...>{});
}

}  // namespace detail
}  // namespace RAJA
