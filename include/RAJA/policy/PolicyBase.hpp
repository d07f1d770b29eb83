// Policy tags and traits (include/RAJA/policy/PolicyBase.hpp:29-254 in the
// reference): every execution policy names its platform and its resource
// type, which the pattern front-ends use to pick a default resource.
#pragma once

#include <type_traits>

#include "RAJA/config.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{

enum class Policy
{
  undefined,
  sequential,
  openmp,
  cdna4
};

namespace policy
{
namespace sequential
{
struct seq_exec
{
  static constexpr Policy policy = Policy::sequential;
  using resource_type = resources::Host;
};
struct seq_reduce
{
};
struct seq_work
{
  using resource_type = resources::Host;
};
}  // namespace sequential

namespace omp
{
struct omp_parallel_for_exec
{
  static constexpr Policy policy = Policy::openmp;
  using resource_type = resources::Host;
};
struct omp_reduce
{
};
struct omp_work
{
  using resource_type = resources::Host;
};
}  // namespace omp
}  // namespace policy

using policy::sequential::seq_exec;
using policy::sequential::seq_reduce;
using policy::sequential::seq_work;
using policy::omp::omp_parallel_for_exec;
using policy::omp::omp_reduce;
using policy::omp::omp_work;

namespace resources
{
template<typename Policy>
struct get_resource
{
  using type = typename Policy::resource_type;
};
}  // namespace resources

namespace type_traits
{
template<typename P>
struct is_sequential_policy : std::integral_constant<bool, P::policy == Policy::sequential>
{
};
template<typename P>
struct is_openmp_policy : std::integral_constant<bool, P::policy == Policy::openmp>
{
};
template<typename P>
struct is_cdna4_policy : std::integral_constant<bool, P::policy == Policy::cdna4>
{
};
template<typename P>
struct is_device_exec_policy : is_cdna4_policy<P>
{
};
}  // namespace type_traits

}  // namespace RAJA
