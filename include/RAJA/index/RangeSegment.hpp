// Iteration spaces of the hot path: contiguous and strided ranges
// (include/RAJA/index/RangeSegment.hpp:97-500 in the reference) and index
// lists (include/RAJA/index/ListSegment.hpp:85-270).  A segment is a
// trivially copyable value handed to kernels; element k is begin()[k].
#pragma once

#include <cstring>
#include <type_traits>

#include "RAJA/config.hpp"

namespace RAJA
{

template<typename T>
struct RangeIterator
{
  T value;
  RAJA_HOST_DEVICE constexpr T operator*() const { return value; }
  RAJA_HOST_DEVICE constexpr T operator[](std::ptrdiff_t k) const { return value + T(k); }
  RAJA_HOST_DEVICE RangeIterator& operator++() { ++value; return *this; }
  RAJA_HOST_DEVICE constexpr std::ptrdiff_t operator-(const RangeIterator& o) const
  {
    return std::ptrdiff_t(value) - std::ptrdiff_t(o.value);
  }
  RAJA_HOST_DEVICE constexpr bool operator!=(const RangeIterator& o) const { return value != o.value; }
};

template<typename StorageT, typename DiffT = std::ptrdiff_t>
struct TypedRangeSegment
{
  using value_type = StorageT;
  using iterator = RangeIterator<StorageT>;
  StorageT m_begin, m_end;
  RAJA_HOST_DEVICE constexpr TypedRangeSegment(StorageT b, StorageT e)
      : m_begin(b), m_end(e < b ? b : e)
  {}
  RAJA_HOST_DEVICE constexpr iterator begin() const { return iterator{m_begin}; }
  RAJA_HOST_DEVICE constexpr iterator end() const { return iterator{m_end}; }
  RAJA_HOST_DEVICE constexpr DiffT size() const { return DiffT(m_end) - DiffT(m_begin); }
  RAJA_HOST_DEVICE constexpr StorageT at(DiffT k) const { return m_begin + StorageT(k); }
  // Elements [begin, begin + length) of this range, clipped to its end.
  RAJA_HOST_DEVICE constexpr TypedRangeSegment slice(DiffT begin, DiffT length) const
  {
    return TypedRangeSegment(clip(begin), clip(begin + length));
  }
  static constexpr bool contiguous = true;

private:
  RAJA_HOST_DEVICE constexpr StorageT clip(DiffT k) const
  {
    return (k <= 0) ? m_begin : (k >= size() ? m_end : m_begin + StorageT(k));
  }
};

template<typename StorageT>
struct StrideIterator
{
  StorageT value, stride;
  RAJA_HOST_DEVICE constexpr StorageT operator*() const { return value; }
  RAJA_HOST_DEVICE constexpr StorageT operator[](std::ptrdiff_t k) const
  {
    return value + StorageT(k) * stride;
  }
};

template<typename StorageT, typename DiffT = std::ptrdiff_t>
struct TypedRangeStrideSegment
{
  using value_type = StorageT;
  using iterator = StrideIterator<StorageT>;
  StorageT m_begin, m_stride;
  DiffT m_size;
  RAJA_HOST_DEVICE constexpr TypedRangeStrideSegment(StorageT b, StorageT e, StorageT stride)
      : m_begin(b), m_stride(stride),
        m_size((stride > 0) ? ((e > b) ? (DiffT(e) - DiffT(b) + stride - 1) / stride : 0)
                            : ((b > e) ? (DiffT(b) - DiffT(e) - stride - 1) / (-stride) : 0))
  {}
  RAJA_HOST_DEVICE constexpr iterator begin() const { return iterator{m_begin, m_stride}; }
  RAJA_HOST_DEVICE constexpr DiffT size() const { return m_size; }
  RAJA_HOST_DEVICE constexpr StorageT at(DiffT k) const { return m_begin + StorageT(k) * m_stride; }
  // Elements [begin, begin + length) of this range, clipped to its end.
  RAJA_HOST_DEVICE TypedRangeStrideSegment slice(DiffT begin, DiffT length) const
  {
    TypedRangeStrideSegment s(*this);
    const DiffT b = begin < 0 ? 0 : (begin > m_size ? m_size : begin);
    const DiffT e = (length > m_size - b) ? m_size : b + (length < 0 ? 0 : length);
    s.m_begin = m_begin + StorageT(b) * m_stride;
    s.m_size = e - b;
    return s;
  }
  static constexpr bool contiguous = false;
};

using RangeSegment = TypedRangeSegment<Index_type>;
using RangeStrideSegment = TypedRangeStrideSegment<Index_type>;

}  // namespace RAJA
