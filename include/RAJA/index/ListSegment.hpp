// TypedListSegment: an index list owned in memory of a resource
// (include/RAJA/index/ListSegment.hpp:85-270 in the reference).  The list
// is copied into the resource's memory at construction (device memory for a
// Cdna4 resource); copies are non-owning views, as kernel arguments are.
#pragma once

#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{

enum IndexOwnership
{
  Unowned,
  Owned
};

template<typename StorageT>
class TypedListSegment
{
public:
  using value_type = StorageT;
  using iterator = const StorageT*;

  template<typename Res>
  TypedListSegment(const StorageT* values, Index_type len, Res res, IndexOwnership own = Owned)
  {
    init(values, len, res, own);
  }

  template<typename Res>
  TypedListSegment(const std::vector<StorageT>& v, Res res)
  {
    init(v.data(), (Index_type)v.size(), res, Owned);
  }

  RAJA_HOST_DEVICE TypedListSegment(const TypedListSegment& o)
      : m_data(o.m_data), m_size(o.m_size), m_owned(false), m_free(nullptr), m_ctx(nullptr)
  {}

  TypedListSegment& operator=(const TypedListSegment&) = delete;

  RAJA_HOST_DEVICE ~TypedListSegment()
  {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (m_owned && m_free) m_free(m_data, m_ctx);
#endif
  }

  RAJA_HOST_DEVICE Index_type size() const { return m_size; }
  RAJA_HOST_DEVICE StorageT at(Index_type k) const { return m_data[k]; }
  RAJA_HOST_DEVICE iterator begin() const { return m_data; }
  RAJA_HOST_DEVICE iterator end() const { return m_data + m_size; }
  RAJA_HOST_DEVICE const StorageT* data() const { return m_data; }
  // Non-owning view of entries [begin, begin + length), clipped to the end.
  RAJA_HOST_DEVICE TypedListSegment slice(Index_type begin, Index_type length) const
  {
    TypedListSegment s(*this);
    const Index_type b = begin < 0 ? 0 : (begin > m_size ? m_size : begin);
    s.m_data = m_data + b;
    s.m_size = (length > m_size - b) ? m_size - b : (length < 0 ? 0 : length);
    return s;
  }

private:
  static void free_host(StorageT* p, void*) { std::free(p); }
#if defined(RAJA_ENABLE_CDNA4)
  static void free_cdna4(StorageT* p, void* stream)
  {
    resources::Cdna4 r((hipStream_t)stream);
    r.deallocate(p);
  }
  void init(const StorageT* values, Index_type len, resources::Cdna4 res, IndexOwnership own)
  {
    m_size = len;
    if (own == Unowned)
    {
      m_data = const_cast<StorageT*>(values);
      return;
    }
    m_data = res.allocate<StorageT>((size_t)len);
    res.memcpy(m_data, values, sizeof(StorageT) * (size_t)len);
    res.wait();
    m_owned = true;
    m_free = &free_cdna4;
    m_ctx = (void*)res.get_stream();
  }
#endif
  void init(const StorageT* values, Index_type len, resources::Host res, IndexOwnership own)
  {
    m_size = len;
    if (own == Unowned)
    {
      m_data = const_cast<StorageT*>(values);
      return;
    }
    m_data = res.allocate<StorageT>((size_t)len);
    std::memcpy(m_data, values, sizeof(StorageT) * (size_t)len);
    m_owned = true;
    m_free = &free_host;
  }

  StorageT* m_data = nullptr;
  Index_type m_size = 0;
  bool m_owned = false;
  void (*m_free)(StorageT*, void*) = nullptr;
  void* m_ctx = nullptr;
};

using ListSegment = TypedListSegment<Index_type>;

}  // namespace RAJA
