// RAJA::operators subset (include/RAJA/util/Operators.hpp:248-652 in the
// reference): identities, binary reduction operators and comparators.
// minimum keeps the left operand on ties ((rhs < lhs) ? rhs : lhs) and maximum
// likewise ((lhs < rhs) ? rhs : lhs), so a left fold returns the first
// extremum.  limits<double> is +-DBL_MAX, not +-inf.
#pragma once

#include <cfloat>
#include <climits>
#include <type_traits>

#include "RAJA/config.hpp"

namespace RAJA
{
namespace operators
{

template<typename T, typename Enable = void>
struct limits
{
  RAJA_HOST_DEVICE static constexpr T min() { return T::min_value(); }
  RAJA_HOST_DEVICE static constexpr T max() { return T::max_value(); }
};
#define RAJA_CDNA4_LIMITS(T, LO, HI)                                            \
  template<>                                                                   \
  struct limits<T>                                                             \
  {                                                                            \
    RAJA_HOST_DEVICE static constexpr T min() { return LO; }                   \
    RAJA_HOST_DEVICE static constexpr T max() { return HI; }                   \
  };
RAJA_CDNA4_LIMITS(char, CHAR_MIN, CHAR_MAX)
RAJA_CDNA4_LIMITS(signed char, SCHAR_MIN, SCHAR_MAX)
RAJA_CDNA4_LIMITS(unsigned char, 0, UCHAR_MAX)
RAJA_CDNA4_LIMITS(short, SHRT_MIN, SHRT_MAX)
RAJA_CDNA4_LIMITS(unsigned short, 0, USHRT_MAX)
RAJA_CDNA4_LIMITS(int, INT_MIN, INT_MAX)
RAJA_CDNA4_LIMITS(unsigned int, 0u, UINT_MAX)
RAJA_CDNA4_LIMITS(long, LONG_MIN, LONG_MAX)
RAJA_CDNA4_LIMITS(unsigned long, 0ul, ULONG_MAX)
RAJA_CDNA4_LIMITS(long long, LLONG_MIN, LLONG_MAX)
RAJA_CDNA4_LIMITS(unsigned long long, 0ull, ULLONG_MAX)
RAJA_CDNA4_LIMITS(float, -FLT_MAX, FLT_MAX)
RAJA_CDNA4_LIMITS(double, -DBL_MAX, DBL_MAX)
#undef RAJA_CDNA4_LIMITS

template<typename Ret, typename Arg1 = Ret, typename Arg2 = Arg1>
struct plus
{
  RAJA_HOST_DEVICE constexpr Ret operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return Ret(lhs) + rhs;
  }
  RAJA_HOST_DEVICE static constexpr Ret identity() { return Ret(0); }
};

template<typename Ret, typename Arg1 = Ret, typename Arg2 = Arg1>
struct minimum
{
  RAJA_HOST_DEVICE constexpr Ret operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return (rhs < lhs) ? Ret(rhs) : Ret(lhs);
  }
  RAJA_HOST_DEVICE static constexpr Ret identity() { return limits<Ret>::max(); }
};

template<typename Ret, typename Arg1 = Ret, typename Arg2 = Arg1>
struct maximum
{
  RAJA_HOST_DEVICE constexpr Ret operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return (lhs < rhs) ? Ret(rhs) : Ret(lhs);
  }
  RAJA_HOST_DEVICE static constexpr Ret identity() { return limits<Ret>::min(); }
};

template<typename Ret, typename Arg1 = Ret, typename Arg2 = Arg1>
struct bit_or
{
  RAJA_HOST_DEVICE constexpr Ret operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return Ret(lhs | rhs);
  }
  RAJA_HOST_DEVICE static constexpr Ret identity() { return Ret(0); }
};

template<typename Ret, typename Arg1 = Ret, typename Arg2 = Arg1>
struct bit_and
{
  RAJA_HOST_DEVICE constexpr Ret operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return Ret(lhs & rhs);
  }
  RAJA_HOST_DEVICE static constexpr Ret identity() { return Ret(~Ret(0)); }
};

template<typename Arg1, typename Arg2 = Arg1>
struct less
{
  RAJA_HOST_DEVICE constexpr bool operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return lhs < rhs;
  }
};

template<typename Arg1, typename Arg2 = Arg1>
struct greater
{
  RAJA_HOST_DEVICE constexpr bool operator()(const Arg1& lhs, const Arg2& rhs) const
  {
    return rhs < lhs;
  }
};

template<typename Op>
struct is_fp_associative : std::false_type
{
};

}  // namespace operators
}  // namespace RAJA
