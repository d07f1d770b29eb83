// cdna4 scans and sorts: the RAJA::impl::scan / impl::sort customisation
// points for cdna4_exec, forwarding pointer spans to libraja_cdna4.so.
//
// Reference counterparts: include/RAJA/policy/hip/scan.hpp:51-274 and
// include/RAJA/policy/hip/sort.hpp:77-689 (which call rocPRIM).  Like the
// reference, the GPU path accepts pointers to arithmetic types with
// operators::plus/minimum/maximum (scans) and operators::less/greater
// (sorts) and rejects anything else at compile time.  Calls are
// asynchronous on the resource's stream; non-async policies wait.
#pragma once

#include <type_traits>

#include "RAJA/policy/cdna4/forall.hpp"
#include "RAJA/util/Operators.hpp"
#include "raja_cdna4.h"

namespace RAJA
{
namespace policy
{
namespace cdna4
{
namespace detail
{

template<typename T>
struct dtype_of;
template<>
struct dtype_of<int32_t>
{
  static constexpr int value = RAJA_CDNA4_I32;
};
template<>
struct dtype_of<int64_t>
{
  static constexpr int value = RAJA_CDNA4_I64;
};
template<>
struct dtype_of<uint32_t>
{
  static constexpr int value = RAJA_CDNA4_U32;
};
template<>
struct dtype_of<uint64_t>
{
  static constexpr int value = RAJA_CDNA4_U64;
};
template<>
struct dtype_of<float>
{
  static constexpr int value = RAJA_CDNA4_F32;
};
template<>
struct dtype_of<double>
{
  static constexpr int value = RAJA_CDNA4_F64;
};
// long long / unsigned long long alias the 64-bit codes on LP64
template<>
struct dtype_of<long long>
{
  static constexpr int value = RAJA_CDNA4_I64;
};
template<>
struct dtype_of<unsigned long long>
{
  static constexpr int value = RAJA_CDNA4_U64;
};

template<typename Fn>
struct op_of
{
  static constexpr int value = -1;
};
template<typename R, typename A, typename B>
struct op_of<operators::plus<R, A, B>>
{
  static constexpr int value = RAJA_CDNA4_OP_PLUS;
};
template<typename R, typename A, typename B>
struct op_of<operators::minimum<R, A, B>>
{
  static constexpr int value = RAJA_CDNA4_OP_MIN;
};
template<typename R, typename A, typename B>
struct op_of<operators::maximum<R, A, B>>
{
  static constexpr int value = RAJA_CDNA4_OP_MAX;
};

template<typename Cmp>
struct order_of
{
  static constexpr int value = -1;
};
template<typename A, typename B>
struct order_of<operators::less<A, B>>
{
  static constexpr int value = 0;
};
template<typename A, typename B>
struct order_of<operators::greater<A, B>>
{
  static constexpr int value = 1;
};

template<bool Async>
resources::EventProxy<resources::Cdna4> finish(resources::Cdna4 res)
{
  if (!Async) res.wait();
  return resources::EventProxy<resources::Cdna4>(res);
}

}  // namespace detail
}  // namespace cdna4
}  // namespace policy

namespace impl
{
namespace scan
{

template<int B, bool Async, int G, typename T, typename BinFn>
resources::EventProxy<resources::Cdna4> inclusive(resources::Cdna4 res,
                                                  const policy::cdna4::cdna4_exec_t<B, Async, G>&,
                                                  const T* begin, const T* end, T* out, BinFn)
{
  namespace d = policy::cdna4::detail;
  static_assert(std::is_arithmetic<T>::value, "cdna4 scans need arithmetic element types");
  static_assert(d::op_of<BinFn>::value >= 0, "cdna4 scans support plus, minimum, maximum");
  resources::detail::check_lib(
      raja_cdna4_inclusive_scan(d::dtype_of<T>::value, d::op_of<BinFn>::value, begin, out,
                                end - begin, res.get_stream()),
      "inclusive_scan");
  return d::finish<Async>(res);
}

template<int B, bool Async, int G, typename T, typename BinFn, typename V>
resources::EventProxy<resources::Cdna4> exclusive(resources::Cdna4 res,
                                                  const policy::cdna4::cdna4_exec_t<B, Async, G>&,
                                                  const T* begin, const T* end, T* out, BinFn,
                                                  V v)
{
  namespace d = policy::cdna4::detail;
  static_assert(std::is_arithmetic<T>::value, "cdna4 scans need arithmetic element types");
  static_assert(d::op_of<BinFn>::value >= 0, "cdna4 scans support plus, minimum, maximum");
  const T init = static_cast<T>(v);
  resources::detail::check_lib(
      raja_cdna4_exclusive_scan(d::dtype_of<T>::value, d::op_of<BinFn>::value, begin, out,
                                end - begin, &init, res.get_stream()),
      "exclusive_scan");
  return d::finish<Async>(res);
}

template<int B, bool Async, int G, typename T, typename BinFn>
resources::EventProxy<resources::Cdna4> inclusive_inplace(
    resources::Cdna4 res, const policy::cdna4::cdna4_exec_t<B, Async, G>& p, T* begin, T* end,
    BinFn f)
{
  return inclusive(res, p, begin, end, begin, f);
}

template<int B, bool Async, int G, typename T, typename BinFn, typename V>
resources::EventProxy<resources::Cdna4> exclusive_inplace(
    resources::Cdna4 res, const policy::cdna4::cdna4_exec_t<B, Async, G>& p, T* begin, T* end,
    BinFn f, V v)
{
  return exclusive(res, p, begin, end, begin, f, v);
}

}  // namespace scan

namespace sort
{

template<int B, bool Async, int G, typename K, typename Compare>
resources::EventProxy<resources::Cdna4> stable(resources::Cdna4 res,
                                               const policy::cdna4::cdna4_exec_t<B, Async, G>&,
                                               K* begin, K* end, Compare)
{
  namespace d = policy::cdna4::detail;
  static_assert(std::is_arithmetic<K>::value, "cdna4 sorts need arithmetic keys");
  static_assert(d::order_of<Compare>::value >= 0, "cdna4 sorts support less and greater");
  resources::detail::check_lib(raja_cdna4_sort_pairs(d::dtype_of<K>::value, begin, nullptr, 0,
                                                     end - begin, d::order_of<Compare>::value,
                                                     res.get_stream()),
                               "sort");
  return d::finish<Async>(res);
}

template<int B, bool Async, int G, typename K, typename Compare>
resources::EventProxy<resources::Cdna4> unstable(resources::Cdna4 res,
                                                 const policy::cdna4::cdna4_exec_t<B, Async, G>& p,
                                                 K* begin, K* end, Compare c)
{
  return stable(res, p, begin, end, c);  // the radix sort is stable
}

template<int B, bool Async, int G, typename K, typename V, typename Compare>
resources::EventProxy<resources::Cdna4> stable_pairs(
    resources::Cdna4 res, const policy::cdna4::cdna4_exec_t<B, Async, G>&, K* kb, K* ke, V* vb,
    Compare)
{
  namespace d = policy::cdna4::detail;
  static_assert(std::is_arithmetic<K>::value, "cdna4 sorts need arithmetic keys");
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "cdna4 sort_pairs values are 4 or 8 bytes");
  static_assert(d::order_of<Compare>::value >= 0, "cdna4 sorts support less and greater");
  resources::detail::check_lib(raja_cdna4_sort_pairs(d::dtype_of<K>::value, kb, vb, (int)sizeof(V),
                                                     ke - kb, d::order_of<Compare>::value,
                                                     res.get_stream()),
                               "sort_pairs");
  return d::finish<Async>(res);
}

template<int B, bool Async, int G, typename K, typename V, typename Compare>
resources::EventProxy<resources::Cdna4> unstable_pairs(
    resources::Cdna4 res, const policy::cdna4::cdna4_exec_t<B, Async, G>& p, K* kb, K* ke, V* vb,
    Compare c)
{
  return stable_pairs(res, p, kb, ke, vb, c);
}

}  // namespace sort
}  // namespace impl
}  // namespace RAJA
