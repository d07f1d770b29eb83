// Forall parameters: RAJA::expt::Reduce / ReduceLoc, ValLoc, ValOp, ValLocOp
// and RAJA::Name, with the interfaces of the reference's
// include/RAJA/pattern/params/params_base.hpp:9-253,
// include/RAJA/pattern/params/reducer.hpp:68-248,
// include/RAJA/pattern/params/kernel_name.hpp and
// include/RAJA/pattern/params/forall.hpp:125-353,599-608:
//   * forall's trailing arguments before the body form a ForallParamPack;
//   * a backend hooks into them through three customization points found by
//     argument-dependent lookup in RAJA::expt::detail --
//       param_init(pol, param, info...)      before the launch,
//       param_combine(pol, param)            at the end of each thread's run
//                                            (device side for GPU policies),
//       param_combine(pol, out, in)          folding two thread copies (host
//                                            policies),
//       param_resolve(pol, param, info...)   after the launch --
//     called for every parameter, in order, by ParamMultiplexer;
//   * the loop body receives one ValOp& per Reduce argument (Name arguments
//     add none), after its index, in declaration order (expt::invoke_body);
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

namespace expt
{
namespace detail
{
// Base of every forall parameter type; its namespace is where the
// customization points live, so ADL finds them for every parameter.
struct ForallParamBase
{
};
}  // namespace detail
}  // namespace expt

namespace detail
{
// Kernel name for tools (include/RAJA/pattern/params/kernel_name.hpp).
struct Name : public expt::detail::ForallParamBase
{
  RAJA_HOST_DEVICE Name() {}
  explicit Name(const char* name_in) : name(name_in) {}
  const char* name = nullptr;
};
}  // namespace detail

inline detail::Name Name(const char* n) { return detail::Name(n); }

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

// One Reduce(&target) argument: Reducer<Op<T,T,T>, T, ValOp<T, Op>>.  The
// body's ValOp lives in m_valop; block_values/nblocks are backend storage
// (cdna4: one value per block of the launch in pinned host memory).
template<typename Op, typename T, typename VOp>
struct Reducer : public ForallParamBase
{
  using op = Op;
  using value_type = T;
  using valop_type = VOp;

  Reducer() = default;
  RAJA_HOST_DEVICE explicit Reducer(value_type* target_in) : target(target_in) {}

  VOp m_valop = VOp{};
  value_type* target = nullptr;

  RAJA_HOST_DEVICE void combineTarget(value_type in)
  {
    const value_type t = op{}(*target, in);
    *target = t;
  }
  RAJA_HOST_DEVICE value_type& getVal() { return m_valop.val; }
  RAJA_HOST_DEVICE VOp* get_lambda_arg() { return &m_valop; }
  static constexpr size_t num_lambda_args = 1;

  value_type* block_values = nullptr;
  int nblocks = 0;
};

// A ValLoc reduction written back to a user ValLoc or to a (value, index)
// pair of targets (ReduceLoc).
template<typename T, typename I, template<typename, typename, typename> class Op>
struct Reducer<Op<ValLoc<T, I>, ValLoc<T, I>, ValLoc<T, I>>, ValLoc<T, I>, ValOp<ValLoc<T, I>, Op>>
    : public ForallParamBase
{
  using target_value_type = T;
  using target_index_type = I;
  using value_type = ValLoc<T, I>;
  using op = Op<value_type, value_type, value_type>;
  using valop_type = ValOp<value_type, Op>;

  Reducer() = default;
  RAJA_HOST_DEVICE explicit Reducer(value_type* target_in)
      : target_value(&target_in->val), target_index(&target_in->loc)
  {
  }
  RAJA_HOST_DEVICE Reducer(target_value_type* data_in, target_index_type* index_in)
      : target_value(data_in), target_index(index_in)
  {
  }

  valop_type m_valop = valop_type{};
  target_value_type* target_value = nullptr;
  target_index_type* target_index = nullptr;

  // operators on ValLoc compare values only and keep the left (target) on ties
  RAJA_HOST_DEVICE void combineTarget(value_type in)
  {
    value_type t(*target_value, *target_index);
    t = op{}(t, in);
    *target_value = t.val;
    *target_index = t.loc;
  }
  RAJA_HOST_DEVICE value_type& getVal() { return m_valop.val; }
  RAJA_HOST_DEVICE valop_type* get_lambda_arg() { return &m_valop; }
  static constexpr size_t num_lambda_args = 1;

  value_type* block_values = nullptr;
  int nblocks = 0;
};

template<typename P>
struct is_instance_of_reducer : std::false_type
{
};
template<typename Op, typename T, typename VOp>
struct is_instance_of_reducer<Reducer<Op, T, VOp>> : std::true_type
{
};

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

// The parallel combine of a Reducer.
template<typename R>
using reducer_tree = typename valop_tree<typename R::valop_type>::type;

}  // namespace detail

// Reduce<Op>(&target): plain value, or a user ValLoc (ValLoc case).
template<template<typename, typename, typename> class Op, typename T>
auto Reduce(T* target)
{
  return detail::Reducer<Op<T, T, T>, T, ValOp<T, Op>>(target);
}

// ReduceLoc<Op>(&value, &index): a ValLoc reduction into two user targets.
template<template<typename, typename, typename> class Op, typename T, typename IndexType>
auto ReduceLoc(T* target, IndexType* index)
{
  using VL = ValLoc<T, IndexType>;
  return detail::Reducer<Op<VL, VL, VL>, VL, ValOp<VL, Op>>(target, index);
}

namespace detail
{
// RAJA::Name needs nothing from any backend: its customization points are
// no-ops for every policy (the tool hooks read it when the forall starts).
template<typename EXEC_POL, typename... Args>
void param_init(EXEC_POL const&, RAJA::detail::Name&, Args&&...)
{
}
template<typename EXEC_POL>
RAJA_HOST_DEVICE void param_combine(EXEC_POL const&, RAJA::detail::Name&)
{
}
template<typename EXEC_POL>
void param_combine(EXEC_POL const&, RAJA::detail::Name&, const RAJA::detail::Name&)
{
}
template<typename EXEC_POL, typename... Args>
void param_resolve(EXEC_POL const&, RAJA::detail::Name&, Args&&...)
{
}
}  // namespace detail

// ---- the parameter pack ------------------------------------------------------------

namespace detail
{
// Indices of the parameters that pass a ValOp to the body (reducers), in
// order.
template<size_t N, typename Acc, typename... Ps>
struct lambda_arg_indices;
template<size_t N, size_t... A>
struct lambda_arg_indices<N, std::index_sequence<A...>>
{
  using type = std::index_sequence<A...>;
};
template<size_t N, size_t... A, typename P, typename... Ps>
struct lambda_arg_indices<N, std::index_sequence<A...>, P, Ps...>
{
  using type = typename std::conditional_t<is_instance_of_reducer<P>::value,
                                           lambda_arg_indices<N + 1, std::index_sequence<A..., N>, Ps...>,
                                           lambda_arg_indices<N + 1, std::index_sequence<A...>, Ps...>>::type;
};
}  // namespace detail

struct ParamMultiplexer;

template<typename... Params>
struct ForallParamPack
{
  friend struct ParamMultiplexer;
  static_assert(std::conjunction<std::is_base_of<detail::ForallParamBase, Params>...>::value,
                "forall parameters must be expt::Reduce / ReduceLoc or RAJA::Name");

  using Base = std::tuple<Params...>;
  Base param_tup;
  static constexpr size_t param_tup_sz = sizeof...(Params);
  using params_seq = std::make_index_sequence<param_tup_sz>;
  using lambda_arg_seq = typename detail::lambda_arg_indices<0, std::index_sequence<>, Params...>::type;

  ForallParamPack() = default;
  explicit ForallParamPack(Base&& t) : param_tup(std::move(t)) {}

private:
  template<typename EXEC_POL, size_t... Seq, typename... Args>
  static void parampack_init(EXEC_POL const& pol, std::index_sequence<Seq...>, ForallParamPack& f,
                             Args&&... args)
  {
    using detail::param_init;
    (void)std::initializer_list<int>{0, (param_init(pol, std::get<Seq>(f.param_tup), args...), 0)...};
  }
  template<typename EXEC_POL, size_t... Seq>
  RAJA_HOST_DEVICE static void parampack_combine(EXEC_POL const& pol, std::index_sequence<Seq...>,
                                                 ForallParamPack& f)
  {
    using detail::param_combine;
    (void)std::initializer_list<int>{0, (param_combine(pol, std::get<Seq>(f.param_tup)), 0)...};
  }
  template<typename EXEC_POL, size_t... Seq>
  static void parampack_combine(EXEC_POL const& pol, std::index_sequence<Seq...>, ForallParamPack& out,
                                const ForallParamPack& in)
  {
    using detail::param_combine;
    (void)std::initializer_list<int>{
        0, (param_combine(pol, std::get<Seq>(out.param_tup), std::get<Seq>(in.param_tup)), 0)...};
  }
  template<typename EXEC_POL, size_t... Seq, typename... Args>
  static void parampack_resolve(EXEC_POL const& pol, std::index_sequence<Seq...>, ForallParamPack& f,
                                Args&&... args)
  {
    using detail::param_resolve;
    (void)std::initializer_list<int>{0, (param_resolve(pol, std::get<Seq>(f.param_tup), args...), 0)...};
  }
};

// How a backend's forall_impl reaches the customization points of every
// parameter (include/RAJA/pattern/params/forall.hpp:247-285 in the reference).
struct ParamMultiplexer
{
  template<typename EXEC_POL, typename... Params, typename... Args,
           typename FP = ForallParamPack<Params...>>
  static void parampack_init(EXEC_POL const& pol, ForallParamPack<Params...>& f_params,
                             Args&&... args)
  {
    FP::parampack_init(pol, typename FP::params_seq(), f_params, std::forward<Args>(args)...);
  }
  template<typename EXEC_POL, typename... Params, typename FP = ForallParamPack<Params...>>
  RAJA_HOST_DEVICE static void parampack_combine(EXEC_POL const& pol,
                                                 ForallParamPack<Params...>& f_params)
  {
    FP::parampack_combine(pol, typename FP::params_seq(), f_params);
  }
  template<typename EXEC_POL, typename... Params, typename FP = ForallParamPack<Params...>>
  static void parampack_combine(EXEC_POL const& pol, ForallParamPack<Params...>& out,
                                const ForallParamPack<Params...>& in)
  {
    FP::parampack_combine(pol, typename FP::params_seq(), out, in);
  }
  template<typename EXEC_POL, typename... Params, typename... Args,
           typename FP = ForallParamPack<Params...>>
  static void parampack_resolve(EXEC_POL const& pol, ForallParamPack<Params...>& f_params,
                                Args&&... args)
  {
    FP::parampack_resolve(pol, typename FP::params_seq(), f_params, std::forward<Args>(args)...);
  }
};

namespace type_traits
{
template<typename T>
struct is_ForallParamPack : std::false_type
{
};
template<typename... Ps>
struct is_ForallParamPack<ForallParamPack<Ps...>> : std::true_type
{
};
template<typename T>
struct is_ForallParamPack_empty : std::true_type
{
};
template<typename P, typename... Ps>
struct is_ForallParamPack_empty<ForallParamPack<P, Ps...>> : std::false_type
{
};
}  // namespace type_traits

namespace detail
{
template<size_t... I, typename Tuple>
auto pack_from(std::index_sequence<I...>, Tuple&& t)
{
  return ForallParamPack<std::decay_t<std::tuple_element_t<I, std::decay_t<Tuple>>>...>(
      std::make_tuple(std::get<I>(std::forward<Tuple>(t))...));
}

template<typename Params, typename Fn, size_t... I, typename... Ts>
RAJA_HOST_DEVICE RAJA_INLINE void invoke_with_order(Params& params, Fn& f, std::index_sequence<I...>,
                                                    Ts&&... extra)
{
  f(std::forward<Ts>(extra)..., std::get<I>(params.param_tup).m_valop...);
}

template<size_t I, size_t N, typename Tuple>
const char* name_in(const Tuple& t)
{
  if constexpr (I == N)
  {
    return nullptr;
  }
  else if constexpr (std::is_same<std::tuple_element_t<I, Tuple>, RAJA::detail::Name>::value)
  {
    return std::get<I>(t).name;
  }
  else
  {
    return name_in<I + 1, N>(t);
  }
}
}  // namespace detail

// (params..., body) -> ForallParamPack of the params (the body is the last
// argument).
template<typename... Args>
auto make_forall_param_pack(Args&&... args)
{
  static_assert(sizeof...(Args) >= 1, "the loop body is the last argument");
  return detail::pack_from(std::make_index_sequence<sizeof...(Args) - 1>{},
                           std::forward_as_tuple(std::forward<Args>(args)...));
}

// The loop body: the last argument.
template<typename... Args>
RAJA_INLINE decltype(auto) get_lambda(Args&&... args)
{
  return std::get<sizeof...(Args) - 1>(std::forward_as_tuple(std::forward<Args>(args)...));
}

// Calls f(extra..., ValOp&...) with the ValOp of every reducer parameter.
template<typename Params, typename Fn, typename... Ts>
RAJA_HOST_DEVICE RAJA_INLINE void invoke_body(Params&& params, Fn&& f, Ts&&... extra)
{
  detail::invoke_with_order(params, f, typename std::decay_t<Params>::lambda_arg_seq(),
                            std::forward<Ts>(extra)...);
}

// The name given by a RAJA::Name parameter, or nullptr.
template<typename... Params>
const char* get_kernel_name(const ForallParamPack<Params...>& f_params)
{
  return detail::name_in<0, sizeof...(Params)>(f_params.param_tup);
}


}  // namespace expt
}  // namespace RAJA
