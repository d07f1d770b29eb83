// Configuration of the cdna4 RAJA front-end subset.
//
// The reference generates RAJA/config.hpp with CMake
// (include/RAJA/config.hpp.in:117-218).  This subset fixes the choices that
// matter to the hot path: 64-bit Index_type (include/RAJA/util/types.hpp:224),
// wave size 64 (config.hpp.in:188), the cdna4 backend when compiled by hipcc
// and the OpenMP backend when compiled with -fopenmp.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RAJA_ENABLE_CDNA4 1
#define RAJA_HOST_DEVICE __host__ __device__
#define RAJA_DEVICE __device__
#else
#define RAJA_HOST_DEVICE
#define RAJA_DEVICE
#endif

#if defined(_OPENMP)
#define RAJA_ENABLE_OPENMP 1
#endif

#define RAJA_INLINE inline __attribute__((always_inline))
#define RAJA_UNUSED_ARG(x)
#define RAJA_UNUSED_VAR(x) (void)(x)

#if defined(__HIP_DEVICE_COMPILE__)
#define RAJA_DEVICE_CODE 1
#endif

namespace RAJA
{
using Index_type = std::ptrdiff_t;
}  // namespace RAJA
