// Span / make_span (include/RAJA/util/Span.hpp:59-214 in the reference):
// the container scans and sorts operate on.
#pragma once

#include <type_traits>

#include "RAJA/config.hpp"

namespace RAJA
{

template<typename IterType, typename IndexType = Index_type>
struct Span
{
  using iterator = IterType;
  using value_type = typename std::remove_reference<decltype(*IterType())>::type;
  IterType m_begin;
  IndexType m_size;
  RAJA_HOST_DEVICE constexpr Span(IterType b, IndexType n) : m_begin(b), m_size(n) {}
  RAJA_HOST_DEVICE constexpr IterType begin() const { return m_begin; }
  RAJA_HOST_DEVICE constexpr IterType end() const { return m_begin + m_size; }
  RAJA_HOST_DEVICE constexpr IndexType size() const { return m_size; }
  RAJA_HOST_DEVICE constexpr bool empty() const { return m_size == 0; }
  RAJA_HOST_DEVICE value_type& operator[](IndexType k) const { return m_begin[k]; }
  RAJA_HOST_DEVICE constexpr IterType data() const { return m_begin; }
};

template<typename IterType, typename IndexType>
RAJA_HOST_DEVICE constexpr Span<IterType, IndexType> make_span(IterType begin, IndexType size)
{
  return Span<IterType, IndexType>(begin, size);
}

}  // namespace RAJA
