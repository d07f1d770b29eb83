// RAJA::synchronize<Policy>() (include/RAJA/pattern/synchronize.hpp:42-46):
// waits for all outstanding work of the policy's platform.
#pragma once

#include "RAJA/config.hpp"
#include "RAJA/policy/PolicyBase.hpp"

namespace RAJA
{

namespace policy
{
namespace sequential
{
inline void synchronize_impl(const seq_exec&) {}
}  // namespace sequential
namespace omp
{
inline void synchronize_impl(const omp_parallel_for_exec&) {}
}  // namespace omp
}  // namespace policy

template<typename Policy>
void synchronize()
{
  synchronize_impl(Policy{});
}

}  // namespace RAJA
