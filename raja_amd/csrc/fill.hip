// Synthetic inputs from a portable counter-based generator: element i of a
// stream with seed s is splitmix64((s << 40) + i) (Steele, Lea and Flood's
// SplitMix64 output function).  Integer outputs are bit-identical to any
// host implementation of the same map, so host checkers can regenerate the
// exact bytes the device used.
#include "runtime.hpp"

namespace
{

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rng_raw(uint64_t seed, uint64_t i)
{
  return mix64((seed << 40) + i + 0x9E3779B97F4A7C15ull);
}

__global__ void fill_kernel(int kind, void* out, int64_t n, uint64_t seed, double a, double b)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
  {
    switch (kind)
    {
      case 0:  // u32, top `a` bits
      {
        const int bits = (int)a;
        static_cast<uint32_t*>(out)[i] = (uint32_t)(rng_raw(seed, (uint64_t)i) >> (64 - bits));
        break;
      }
      case 1:  // i32 uniform in [a, b)
      {
        const uint64_t span = (uint64_t)((int64_t)b - (int64_t)a);
        static_cast<int32_t*>(out)[i] =
            (int32_t)((int64_t)a + (int64_t)(((rng_raw(seed, (uint64_t)i) >> 32) * span) >> 32));
        break;
      }
      case 2:  // i64 uniform in [a, b)
      {
        const uint64_t span = (uint64_t)((int64_t)b - (int64_t)a);
        static_cast<int64_t*>(out)[i] =
            (int64_t)a + (int64_t)(((rng_raw(seed, (uint64_t)i) >> 32) * span) >> 32);
        break;
      }
      case 3:  // f64 U[a, b)
      {
        const double u = (double)(rng_raw(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
        static_cast<double*>(out)[i] = __dadd_rn(a, __dmul_rn(b - a, u));
        break;
      }
      case 4:  // f64 N(0,1), Box-Muller on counters 2k, 2k+1
      {
        const uint64_t k = (uint64_t)i >> 1;
        const double u1 = ((double)(rng_raw(seed, 2 * k) >> 11) + 1.0) * 0x1.0p-53;
        const double u2 = (double)(rng_raw(seed, 2 * k + 1) >> 11) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(u1));
        const double t = 6.283185307179586 * u2;
        static_cast<double*>(out)[i] = (i & 1) ? r * sin(t) : r * cos(t);
        break;
      }
      case 5:  // i32 iota
        static_cast<int32_t*>(out)[i] = (int32_t)((int64_t)a + i);
        break;
      case 6:  // f64 a + i
        static_cast<double*>(out)[i] = a + (double)i;
        break;
      default:  // u32 constant
        static_cast<uint32_t*>(out)[i] = (uint32_t)a;
        break;
    }
  }
}

}  // namespace

extern "C" int raja_cdna4_fill(int kind, void* out, int64_t n, uint64_t seed, double a, double b,
                               void* stream)
{
  if (n < 0 || kind < 0 || kind > 7) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad fill");
  if (kind == 0 && (a < 1 || a > 32)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bits");
  if (n == 0) return RAJA_CDNA4_SUCCESS;
  if (!out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "out is NULL");
  const int64_t want = rc::ceil_div(n, 256);
  const int64_t cap = (int64_t)rc::num_cus() * 16;
  const unsigned grid = (unsigned)(want < cap ? want : cap);
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, kind, out, n,
                     seed, a, b);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}
