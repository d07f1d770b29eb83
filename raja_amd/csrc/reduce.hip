// Deterministic span reductions (ReduceSum/Min/Max/BitOr/BitAnd,
// MinLoc/MaxLoc and the fused Sum+Min+Max+MinLoc of one forall).
//
// Reference: the HIP reducers fold per-thread values with a shuffle tree,
// per-block slots and a last-block-done pass, replicated 32 ways and folded
// again on the host (include/RAJA/policy/hip/reduce.hpp:127-208, 960-1161).
// Here the span is cut into fixed tiles of TILE elements.  Kernel 1 folds
// each tile in a fixed order (4 interleaved accumulators per thread, DPP wave
// tree, waves in order) and stores the tile value; kernel 2, one block, folds
// the tile values in a fixed order and applies the caller's init.  The
// combine tree depends only on n -- not on the grid, the timing or whatever
// else shares the GPU -- and no floating-point atomics are used.  Two
// kernels instead of a last-block-done pass: on MI355X a release fence at
// device scope writes back the whole L2 of the block's XCD, and a
// same-address device atomic per block serialises across the 8 XCDs; the
// kernel boundary gives the visibility for free.
#include "ops.hpp"
#include "runtime.hpp"

namespace
{
using namespace RAJA::cdna4::dev;

constexpr int kBlock = 256;
constexpr int kVecPerChunk = 16;  // 16-byte loads in flight per thread
constexpr int kChunks = 4;        // chunks per tile
constexpr int kFinalBlock = 1024;

template<typename T>
struct VecOf
{
  static constexpr int n = 16 / sizeof(T);
  typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};

template<typename T>
constexpr int64_t tile_elems()
{
  return (int64_t)kBlock * kVecPerChunk * kChunks * VecOf<T>::n;
}

// Fold of one tile by the whole block.  Thread t owns vectors k*B + t of the
// tile (k = 0..63), accumulated into acc[k % 4]; returns the tile value.
template<typename T, typename Op, bool ALIGNED>
__device__ __forceinline__ typename Op::acc_t fold_tile(const T* __restrict__ x, int64_t n,
                                                        int64_t tile,
                                                        typename Op::acc_t* s_scratch)
{
  using acc_t = typename Op::acc_t;
  using vec_t = typename VecOf<T>::type;
  constexpr int VEC = VecOf<T>::n;
  constexpr int64_t TILE = tile_elems<T>();
  const Op op{};
  const int t = (int)threadIdx.x;
  const int64_t base = tile * TILE;
  acc_t acc[4] = {Op::ident(), Op::ident(), Op::ident(), Op::ident()};
  if (ALIGNED && base + TILE <= n)
  {
    for (int c = 0; c < kChunks; ++c)
    {
      vec_t v[kVecPerChunk];
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int k = c * kVecPerChunk + j;
        v[j] = *reinterpret_cast<const vec_t*>(x + base + ((int64_t)k * kBlock + t) * VEC);
      }
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int k = c * kVecPerChunk + j;
        const int64_t i0 = base + ((int64_t)k * kBlock + t) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
          acc[j & 3] = op(acc[j & 3], Op::lift(v[j][e], i0 + e));
        }
      }
    }
  }
  else
  {
    for (int k = 0; k < kVecPerChunk * kChunks; ++k)
    {
      const int64_t i0 = base + ((int64_t)k * kBlock + t) * VEC;
#pragma unroll
      for (int e = 0; e < VEC; ++e)
      {
        if (i0 + e < n)
        {
          acc[k & 3] = op(acc[k & 3], Op::lift(x[i0 + e], i0 + e));
        }
      }
    }
  }
  acc_t a = op(op(acc[0], acc[1]), op(acc[2], acc[3]));
  return block_reduce(a, op, s_scratch);
}

template<typename T, typename Op, bool ALIGNED>
__global__ __launch_bounds__(kBlock) void reduce_tiles_kernel(
    const T* __restrict__ x, int64_t n, int64_t ntiles, typename Op::acc_t* __restrict__ partials)
{
  using acc_t = typename Op::acc_t;
  __shared__ acc_t s_scratch[max_waves];
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
  {
    acc_t v = fold_tile<T, Op, ALIGNED>(x, n, tile, s_scratch);
    if (threadIdx.x == 0)
    {
      partials[tile] = v;
    }
  }
}

// Fixed-order fold of the tile values: thread t takes tiles t, t+1024, ...
// into 8 interleaved accumulators, then the block tree.
template<typename Op>
__global__ __launch_bounds__(kFinalBlock) void reduce_final_kernel(
    const typename Op::acc_t* __restrict__ partials, int64_t ntiles, typename Op::acc_t init,
    typename Op::acc_t* __restrict__ out)
{
  using acc_t = typename Op::acc_t;
  const Op op{};
  __shared__ acc_t s_scratch[max_waves];
  acc_t acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = Op::ident();
  int j = 0;
  for (int64_t k = threadIdx.x; k < ntiles; k += kFinalBlock)
  {
    acc[j] = op(acc[j], partials[k]);
    j = (j + 1) & 7;
  }
  acc_t a = op(op(op(acc[0], acc[1]), op(acc[2], acc[3])),
               op(op(acc[4], acc[5]), op(acc[6], acc[7])));
  a = block_reduce(a, op, s_scratch);
  if (threadIdx.x == 0)
  {
    *out = Op::with_init(init, a);
  }
}

template<typename T, typename Op>
int launch_reduce(const T* x, int64_t n, typename Op::acc_t init,
                  typename Op::acc_t* out, hipStream_t s, const char* name)
{
  using acc_t = typename Op::acc_t;
  if (n < 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "negative length");
  if (!out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "output is NULL");
  if (n > 0 && !x) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "input is NULL");
  rc::StreamCtx* c = nullptr;
  RC_TRY(rc::get_ctx(s, &c));
  const int64_t ntiles = (n == 0) ? 1 : rc::ceil_div(n, tile_elems<T>());
  RC_TRY(rc::ensure_tmp(c, (size_t)ntiles * sizeof(acc_t)));
  const int64_t max_blocks = (int64_t)rc::num_cus() * 8;
  const unsigned grid = (unsigned)(ntiles < max_blocks ? ntiles : max_blocks);
  const bool aligned = (((uintptr_t)x) % 16) == 0;
  acc_t* partials = (acc_t*)c->tmp;
  rc::TimedScope ts(name, s);
  if (aligned)
  {
    hipLaunchKernelGGL((reduce_tiles_kernel<T, Op, true>), dim3(grid), dim3(kBlock), 0, s, x, n,
                       ntiles, partials);
  }
  else
  {
    hipLaunchKernelGGL((reduce_tiles_kernel<T, Op, false>), dim3(grid), dim3(kBlock), 0, s, x,
                       n, ntiles, partials);
  }
  RC_HIP(hipGetLastError());
  hipLaunchKernelGGL((reduce_final_kernel<Op>), dim3(1), dim3(kFinalBlock), 0, s, partials, ntiles,
                     init, out);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}


template<typename T>
int reduce_typed(int op, const void* x, int64_t n, const void* init, void* out, hipStream_t s)
{
  const T iv = *static_cast<const T*>(init);
  switch (op)
  {
    case RAJA_CDNA4_OP_PLUS:
      return launch_reduce<T, rk::OpPlus<T>>((const T*)x, n, iv, (T*)out, s, "reduce_sum");
    case RAJA_CDNA4_OP_MIN:
      return launch_reduce<T, rk::OpMin<T>>((const T*)x, n, iv, (T*)out, s, "reduce_min");
    case RAJA_CDNA4_OP_MAX:
      return launch_reduce<T, rk::OpMax<T>>((const T*)x, n, iv, (T*)out, s, "reduce_max");
    default:
      return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "operator %d not supported here", op);
  }
}

template<typename T>
int reduce_bits(int op, const void* x, int64_t n, const void* init, void* out, hipStream_t s)
{
  const T iv = *static_cast<const T*>(init);
  switch (op)
  {
    case RAJA_CDNA4_OP_BIT_OR:
      return launch_reduce<T, rk::OpBitOr<T>>((const T*)x, n, iv, (T*)out, s, "reduce_bitor");
    case RAJA_CDNA4_OP_BIT_AND:
      return launch_reduce<T, rk::OpBitAnd<T>>((const T*)x, n, iv, (T*)out, s, "reduce_bitand");
    default:
      return reduce_typed<T>(op, x, n, init, out, s);
  }
}

// Loc results land in a device ValLoc; a 1-thread kernel splits it into the
// caller's value and index slots.
template<typename T>
__global__ void split_valloc_kernel(const rk::ValLoc<T>* in, T* val, int64_t* loc)
{
  *val = in->val;
  *loc = in->loc;
}

template<typename T>
int reduce_loc_typed(int is_max, const void* x, int64_t n, const void* init_val, int64_t init_loc,
                     void* out_val, int64_t* out_loc, hipStream_t s)
{
  rc::StreamCtx* c = nullptr;
  RC_TRY(rc::get_ctx(s, &c));
  // The ValLoc result goes to the stream's result slot, then is split.
  rk::ValLoc<T>* my = static_cast<rk::ValLoc<T>*>(c->small);
  const rk::ValLoc<T> iv = rk::make_valloc<T>(*static_cast<const T*>(init_val), init_loc);
  if (is_max)
  {
    RC_TRY((launch_reduce<T, rk::OpMaxLoc<T>>((const T*)x, n, iv, my, s, "reduce_maxloc")));
  }
  else
  {
    RC_TRY((launch_reduce<T, rk::OpMinLoc<T>>((const T*)x, n, iv, my, s, "reduce_minloc")));
  }
  hipLaunchKernelGGL((split_valloc_kernel<T>), dim3(1), dim3(1), 0, s, my, (T*)out_val, out_loc);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}

}  // namespace

extern "C" {

int raja_cdna4_reduce(int dtype, int op, const void* x, int64_t n, const void* init,
                      void* out_dev, void* stream)
{
  if (!init) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "init is NULL");
  hipStream_t s = (hipStream_t)stream;
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return reduce_bits<int32_t>(op, x, n, init, out_dev, s);
    case RAJA_CDNA4_I64: return reduce_bits<int64_t>(op, x, n, init, out_dev, s);
    case RAJA_CDNA4_U32: return reduce_bits<uint32_t>(op, x, n, init, out_dev, s);
    case RAJA_CDNA4_U64: return reduce_bits<uint64_t>(op, x, n, init, out_dev, s);
    case RAJA_CDNA4_F32: return reduce_typed<float>(op, x, n, init, out_dev, s);
    case RAJA_CDNA4_F64: return reduce_typed<double>(op, x, n, init, out_dev, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}

int raja_cdna4_reduce_loc(int dtype, int is_max, const void* x, int64_t n, const void* init_val,
                          int64_t init_loc, void* out_val_dev, int64_t* out_loc_dev, void* stream)
{
  if (!init_val || !out_val_dev || !out_loc_dev)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL argument");
  }
  hipStream_t s = (hipStream_t)stream;
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return reduce_loc_typed<int32_t>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    case RAJA_CDNA4_I64: return reduce_loc_typed<int64_t>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    case RAJA_CDNA4_U32: return reduce_loc_typed<uint32_t>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    case RAJA_CDNA4_U64: return reduce_loc_typed<uint64_t>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    case RAJA_CDNA4_F32: return reduce_loc_typed<float>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    case RAJA_CDNA4_F64: return reduce_loc_typed<double>(is_max, x, n, init_val, init_loc, out_val_dev, out_loc_dev, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}

int raja_cdna4_multi_reduce_f64(const double* x, int64_t n, raja_cdna4_multi_f64* out_dev,
                                void* stream)
{
  static_assert(sizeof(raja_cdna4_multi_f64) == sizeof(rk::Multi), "layout");
  return launch_reduce<double, rk::OpMulti>(x, n, rk::OpMulti::ident(), (rk::Multi*)out_dev,
                                            (hipStream_t)stream, "reduce_multi_f64");
}

}  // extern "C"
