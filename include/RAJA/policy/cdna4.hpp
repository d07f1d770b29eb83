// The cdna4 (MI355X, gfx950) policy family.
#pragma once

#include "RAJA/policy/cdna4/forall.hpp"
#include "RAJA/policy/cdna4/algorithms.hpp"
#include "RAJA/policy/cdna4/WorkGroup.hpp"
#include "RAJA/policy/cdna4/launch.hpp"
#include "RAJA/policy/cdna4/multi_reduce.hpp"

namespace RAJA
{
namespace policy
{
namespace cdna4
{
// synchronize<cdna4_synchronize>() waits for the whole device
// (include/RAJA/policy/hip/synchronize.hpp:40-43).
inline void synchronize_impl(const cdna4_synchronize&)
{
  resources::detail::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
}  // namespace cdna4
}  // namespace policy
}  // namespace RAJA
