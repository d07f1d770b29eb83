// Radix-sort key handling shared by the onesweep and block-local sorts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace
{

constexpr int kRadix = 256;  // 8-bit digits

// Order-preserving map of a key to unsigned bits (operator< order).
template<typename K>
struct KeyTraits;
template<>
struct KeyTraits<uint32_t>
{
  using bits_t = uint32_t;
  __device__ static bits_t to_bits(uint32_t k) { return k; }
};
template<>
struct KeyTraits<int32_t>
{
  using bits_t = uint32_t;
  __device__ static bits_t to_bits(int32_t k) { return (uint32_t)k ^ 0x80000000u; }
};
template<>
struct KeyTraits<float>
{
  using bits_t = uint32_t;
  __device__ static bits_t to_bits(float k)
  {
    uint32_t u = __float_as_uint(k);
    u = (u == 0x80000000u) ? 0u : u;  // -0.0 == +0.0 under operator<: one key
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
};
template<>
struct KeyTraits<uint64_t>
{
  using bits_t = uint64_t;
  __device__ static bits_t to_bits(uint64_t k) { return k; }
};
template<>
struct KeyTraits<int64_t>
{
  using bits_t = uint64_t;
  __device__ static bits_t to_bits(int64_t k) { return (uint64_t)k ^ 0x8000000000000000ull; }
};
template<>
struct KeyTraits<double>
{
  using bits_t = uint64_t;
  __device__ static bits_t to_bits(double k)
  {
    uint64_t u = (uint64_t)__double_as_longlong(k);
    u = (u == 0x8000000000000000ull) ? 0ull : u;  // -0.0 == +0.0 under operator<
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
  }
};

template<typename K, bool DESC>
__device__ __forceinline__ uint32_t digit_of(K k, int shift)
{
  auto b = KeyTraits<K>::to_bits(k);
  if (DESC) b = ~b;
  return (uint32_t)(b >> shift) & (kRadix - 1);
}

}  // namespace
