// Operator and accumulator types shared by the library kernels.
//
// Identities follow RAJA::operators (include/RAJA/util/Operators.hpp:248-330,
// 363-595): plus -> 0, minimum -> limits<T>::max() (DBL_MAX for double, not
// inf), maximum -> limits<T>::min() (-DBL_MAX for double, INT_MIN for int),
// bit_or -> 0, bit_and -> ~0.  Value/index pairs compare lexicographically
// (value, then lower index), so a reduction returns the first extremum in
// index order -- the seq_exec result (include/RAJA/util/Operators.hpp:555-595,
// ties keep the left operand, which is the earlier index).
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include <type_traits>

namespace rk
{

template<typename T>
struct num_limits;
template<>
struct num_limits<int32_t>
{
  __host__ __device__ static constexpr int32_t max() { return INT32_MAX; }
  __host__ __device__ static constexpr int32_t lowest() { return INT32_MIN; }
};
template<>
struct num_limits<int64_t>
{
  __host__ __device__ static constexpr int64_t max() { return INT64_MAX; }
  __host__ __device__ static constexpr int64_t lowest() { return INT64_MIN; }
};
template<>
struct num_limits<uint32_t>
{
  __host__ __device__ static constexpr uint32_t max() { return UINT32_MAX; }
  __host__ __device__ static constexpr uint32_t lowest() { return 0u; }
};
template<>
struct num_limits<uint64_t>
{
  __host__ __device__ static constexpr uint64_t max() { return UINT64_MAX; }
  __host__ __device__ static constexpr uint64_t lowest() { return 0ull; }
};
template<>
struct num_limits<float>
{
  __host__ __device__ static constexpr float max() { return FLT_MAX; }
  __host__ __device__ static constexpr float lowest() { return -FLT_MAX; }
};
template<>
struct num_limits<double>
{
  __host__ __device__ static constexpr double max() { return DBL_MAX; }
  __host__ __device__ static constexpr double lowest() { return -DBL_MAX; }
};

template<typename T>
struct OpPlus
{
  using acc_t = T;
  __host__ __device__ static acc_t ident() { return T(0); }
  __device__ static acc_t lift(T v, int64_t) { return v; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const { return a + b; }
  // Final combine with the caller's init (seq: folded only if not identity).
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r)
  {
    return (r != ident()) ? init + r : init;
  }
};

template<typename T>
struct OpMin
{
  using acc_t = T;
  __host__ __device__ static acc_t ident() { return num_limits<T>::max(); }
  __device__ static acc_t lift(T v, int64_t) { return v; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const { return (b < a) ? b : a; }
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r)
  {
    return (r < init) ? r : init;
  }
};

template<typename T>
struct OpMax
{
  using acc_t = T;
  __host__ __device__ static acc_t ident() { return num_limits<T>::lowest(); }
  __device__ static acc_t lift(T v, int64_t) { return v; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const { return (a < b) ? b : a; }
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r)
  {
    return (init < r) ? r : init;
  }
};

template<typename T>
struct OpBitOr
{
  using acc_t = T;
  __host__ __device__ static acc_t ident() { return T(0); }
  __device__ static acc_t lift(T v, int64_t) { return v; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const { return a | b; }
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r) { return init | r; }
};

template<typename T>
struct OpBitAnd
{
  using acc_t = T;
  __host__ __device__ static acc_t ident() { return (T)~(T)0; }
  __device__ static acc_t lift(T v, int64_t) { return v; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const { return a & b; }
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r) { return init & r; }
};

// Value with index; 16-byte layout for every T so DPP moves whole words.
template<typename T>
struct alignas(8) ValLoc
{
  T val;
  int32_t pad[(sizeof(T) == 4) ? 1 : 2];
  int64_t loc;
};

template<typename T>
__host__ __device__ inline ValLoc<T> make_valloc(T v, int64_t i)
{
  ValLoc<T> r;
  r.val = v;
  r.pad[0] = 0;
  if (sizeof(r.pad) > 4) r.pad[sizeof(r.pad) / 4 - 1] = 0;
  r.loc = i;
  return r;
}

template<typename T>
struct OpMinLoc
{
  using acc_t = ValLoc<T>;
  __host__ __device__ static acc_t ident() { return make_valloc<T>(num_limits<T>::max(), -1); }
  __device__ static acc_t lift(T v, int64_t i) { return make_valloc<T>(v, i); }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const
  {
    return (b.val < a.val || (b.val == a.val && b.loc < a.loc)) ? b : a;
  }
  // Element i folded into a, for elements met in increasing index order
  // (a tie keeps the earlier index, as operator() does).
  __device__ static void accumulate(acc_t& a, T v, int64_t i)
  {
    const bool take = v < a.val;  // selects, not a branch
    a.val = take ? v : a.val;
    a.loc = take ? i : a.loc;
  }
  // The initial pair wins ties (seq folds the loop result in with a strict <).
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r)
  {
    return (r.val < init.val) ? r : init;
  }
};

template<typename T>
struct OpMaxLoc
{
  using acc_t = ValLoc<T>;
  __host__ __device__ static acc_t ident() { return make_valloc<T>(num_limits<T>::lowest(), -1); }
  __device__ static acc_t lift(T v, int64_t i) { return make_valloc<T>(v, i); }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const
  {
    return (a.val < b.val || (b.val == a.val && b.loc < a.loc)) ? b : a;
  }
  __device__ static void accumulate(acc_t& a, T v, int64_t i)
  {
    const bool take = a.val < v;
    a.val = take ? v : a.val;
    a.loc = take ? i : a.loc;
  }
  __host__ __device__ static acc_t with_init(acc_t init, acc_t r)
  {
    return (init.val < r.val) ? r : init;
  }
};

// Sum, min, max and min-with-index of FP64 data in one accumulator.
struct Multi
{
  double sum, mn, mx, lv;
  int64_t ll, pad;
};

struct OpMulti
{
  using acc_t = Multi;
  __host__ __device__ static acc_t ident()
  {
    return Multi{0.0, DBL_MAX, -DBL_MAX, DBL_MAX, -1, 0};
  }
  __device__ static acc_t lift(double v, int64_t i) { return Multi{v, v, v, v, i, 0}; }
  __host__ __device__ acc_t operator()(acc_t a, acc_t b) const
  {
    Multi r;
    r.sum = a.sum + b.sum;
    r.mn = (b.mn < a.mn) ? b.mn : a.mn;
    r.mx = (a.mx < b.mx) ? b.mx : a.mx;
    const bool tb = (b.lv < a.lv) || (b.lv == a.lv && b.ll < a.ll);
    r.lv = tb ? b.lv : a.lv;
    r.ll = tb ? b.ll : a.ll;
    r.pad = 0;
    return r;
  }
  __host__ __device__ static acc_t with_init(acc_t, acc_t r) { return r; }
  // Element i folded into a, for elements met in increasing index order.
  __device__ static void accumulate(acc_t& a, double v, int64_t i)
  {
    a.sum += v;
    a.mn = (v < a.mn) ? v : a.mn;
    a.mx = (a.mx < v) ? v : a.mx;
    const bool take = v < a.lv;  // selects, not a branch
    a.lv = take ? v : a.lv;
    a.ll = take ? i : a.ll;
  }
};

// enum raja_cdna4_dtype code of an element type.
template<typename T>
constexpr int dtype_code()
{
  if constexpr (std::is_same_v<T, int32_t>) return 0;
  else if constexpr (std::is_same_v<T, int64_t>) return 1;
  else if constexpr (std::is_same_v<T, uint32_t>) return 2;
  else if constexpr (std::is_same_v<T, uint64_t>) return 3;
  else if constexpr (std::is_same_v<T, float>) return 4;
  else return 5;
}

}  // namespace rk
