// Compile-time check of the drop-in boundary: the cdna4 policies provide the
// customization points with the argument shapes the reference's pattern
// layer calls them with, and they instantiate.
//
//   * expt reducers: param_init(pol, Reducer<Op,T,VOp>&, info&),
//     param_combine(pol, Reducer&) (device side, at kernel end) and
//     param_resolve(pol, Reducer&, info&), reached through
//     expt::ParamMultiplexer (include/RAJA/pattern/params/forall.hpp:247-285,
//     include/RAJA/policy/hip/params/reduce.hpp:19-60);
//   * forall_impl(res, pol, iterable, body, ForallParamPack) with the pack by
//     value (include/RAJA/policy/hip/forall.hpp:543-560);
//   * detail::WorkRunner<Exec, Order, Dispatch, Allocator, Index, Args...>
//     with enqueue(storage, iterable, body), run(storage, res, args...) ->
//     per_run_storage, clear(), dispatcher_type and resource_type
//     (include/RAJA/policy/hip/WorkGroup/WorkRunner.hpp:283-387).
//
// Built with hipcc for gfx950 (host and device passes); nothing runs.
#include <memory>
#include <type_traits>
#include <utility>

#include "RAJA/RAJA.hpp"

namespace
{
using Pol = RAJA::cdna4_exec<256>;
using LPol = RAJA::policy::cdna4::cdna4_launch_t<false, 256>;
using Info = RAJA::cdna4::detail::cdna4Info;
using Sum = RAJA::expt::detail::Reducer<RAJA::operators::plus<double, double, double>, double,
                                        RAJA::expt::ValOp<double, RAJA::operators::plus>>;
using VL = RAJA::expt::ValLoc<double, RAJA::Index_type>;
using MinLoc = RAJA::expt::detail::Reducer<RAJA::operators::minimum<VL, VL, VL>, VL,
                                           RAJA::expt::ValOp<VL, RAJA::operators::minimum>>;

// Reduce / ReduceLoc build the reference's Reducer<Op<T,T,T>, T, ValOp<T,Op>>.
static_assert(std::is_same<decltype(RAJA::expt::Reduce<RAJA::operators::plus>((double*)nullptr)),
                           Sum>::value,
              "Reduce<plus>(double*)");
static_assert(std::is_same<decltype(RAJA::expt::ReduceLoc<RAJA::operators::minimum>(
                               (double*)nullptr, (RAJA::Index_type*)nullptr)),
                           MinLoc>::value,
              "ReduceLoc<minimum>(double*, Index_type*)");
static_assert(std::is_same<decltype(RAJA::expt::Reduce<RAJA::operators::minimum>((VL*)nullptr)),
                           MinLoc>::value,
              "Reduce<minimum>(ValLoc*)");

// The customization points, found by argument-dependent lookup as the
// multiplexer finds them.
template<typename P, typename R>
using init_t = decltype(param_init(std::declval<const P&>(), std::declval<R&>(), std::declval<const Info&>()));
template<typename P, typename R>
using combine_t = decltype(param_combine(std::declval<const P&>(), std::declval<R&>()));
template<typename P, typename R>
using resolve_t = decltype(param_resolve(std::declval<const P&>(), std::declval<R&>(), std::declval<const Info&>()));

static_assert(std::is_void<init_t<Pol, Sum>>::value && std::is_void<combine_t<Pol, Sum>>::value &&
                  std::is_void<resolve_t<Pol, Sum>>::value,
              "cdna4 forall: param_init/combine/resolve for Reducer");
static_assert(std::is_void<init_t<Pol, MinLoc>>::value && std::is_void<combine_t<Pol, MinLoc>>::value &&
                  std::is_void<resolve_t<Pol, MinLoc>>::value,
              "cdna4 forall: param_init/combine/resolve for the ValLoc Reducer");
static_assert(std::is_void<init_t<LPol, Sum>>::value && std::is_void<resolve_t<LPol, Sum>>::value,
              "cdna4 launch: param_init/resolve for Reducer");
static_assert(std::is_void<init_t<Pol, RAJA::detail::Name>>::value &&
                  std::is_void<combine_t<Pol, RAJA::detail::Name>>::value &&
                  std::is_void<resolve_t<Pol, RAJA::detail::Name>>::value,
              "Name parameters");

// The pack: reducers pass one ValOp& each to the body, Name passes none.
struct Body
{
  RAJA_HOST_DEVICE void operator()(RAJA::Index_type, RAJA::expt::ValOp<double, RAJA::operators::plus>&,
                                   RAJA::expt::ValOp<VL, RAJA::operators::minimum>&) const
  {
  }
};
using FP = decltype(RAJA::expt::make_forall_param_pack(std::declval<Sum>(), RAJA::Name("n"),
                                                       std::declval<MinLoc>(), Body{}));
static_assert(std::is_same<FP, RAJA::expt::ForallParamPack<Sum, RAJA::detail::Name, MinLoc>>::value,
              "ForallParamPack of the parameters, the body stripped");
static_assert(std::is_same<FP::lambda_arg_seq, std::index_sequence<0, 2>>::value,
              "lambda arguments come from the reducers only");
static_assert(RAJA::expt::type_traits::is_ForallParamPack<FP>::value &&
                  !RAJA::expt::type_traits::is_ForallParamPack_empty<FP>::value,
              "pack traits");

// forall_impl takes the pack by value and returns an EventProxy.
using FI = decltype(RAJA::policy::cdna4::forall_impl(std::declval<RAJA::resources::Cdna4>(), Pol{},
                                                     std::declval<RAJA::TypedRangeSegment<RAJA::Index_type>>(),
                                                     Body{}, std::declval<FP>()));
static_assert(std::is_same<FI, RAJA::resources::EventProxy<RAJA::resources::Cdna4>>::value,
              "forall_impl(res, pol, iter, body, ForallParam)");

// WorkRunner interface.
struct LoopBody
{
  double* p;
  RAJA_HOST_DEVICE void operator()(RAJA::Index_type i, int v) const { p[i] = v; }
};
template<typename Order, typename Dispatch>
using WR = RAJA::detail::WorkRunner<RAJA::cdna4_work<256>, Order, Dispatch, std::allocator<char>,
                                    RAJA::Index_type, int>;
using WS = RAJA::detail::workgroup::LoopStorage<RAJA::Index_type, int>;

template<typename R>
constexpr bool runner_shape()
{
  static_assert(std::is_same<typename R::resource_type, RAJA::resources::Cdna4>::value, "resource_type");
  static_assert(std::is_same<decltype(std::declval<const R&>().run(std::declval<const WS&>(),
                                                                   std::declval<RAJA::resources::Cdna4>(), 1)),
                             typename R::per_run_storage>::value,
                "run(storage, res, args...) -> per_run_storage");
  static_assert(std::is_void<decltype(std::declval<R&>().enqueue(
                    std::declval<WS&>(), std::declval<RAJA::TypedRangeSegment<RAJA::Index_type>>(),
                    std::declval<LoopBody>()))>::value,
                "enqueue(storage, iter, body)");
  static_assert(std::is_void<decltype(std::declval<R&>().clear())>::value, "clear()");
  static_assert(sizeof(typename R::dispatcher_type) > 0, "dispatcher_type");
  return true;
}
static_assert(runner_shape<WR<RAJA::ordered, RAJA::indirect_function_call_dispatch>>(), "ordered");
static_assert(runner_shape<WR<RAJA::unordered_cdna4_loop_tiles,
                              RAJA::direct_dispatch<camp::list<RAJA::TypedRangeSegment<RAJA::Index_type>, LoopBody>>>>(),
              "unordered, direct dispatch");

// Instantiate the definitions (not only the declarations).
[[maybe_unused]] void instantiate(RAJA::resources::Cdna4 res, double* p)
{
  FP fp = RAJA::expt::make_forall_param_pack(Sum(p), RAJA::Name("n"), MinLoc(p, (RAJA::Index_type*)nullptr),
                                             Body{});
  Info info(res, dim3(1), dim3(256));
  RAJA::expt::ParamMultiplexer::parampack_init(Pol{}, fp, info);
  RAJA::expt::ParamMultiplexer::parampack_resolve(Pol{}, fp, info);
  RAJA::policy::cdna4::forall_impl(res, Pol{}, RAJA::TypedRangeSegment<RAJA::Index_type>(0, 4), Body{}, fp);
  WS storage;
  WR<RAJA::ordered, RAJA::indirect_function_call_dispatch> runner;
  runner.enqueue(storage, RAJA::TypedRangeSegment<RAJA::Index_type>(0, 4), LoopBody{p});
  (void)runner.run(storage, res, 3);
  runner.clear();
}
}  // namespace

int main() { return 0; }
