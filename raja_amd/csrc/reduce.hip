// Deterministic span reductions (ReduceSum/Min/Max/BitOr/BitAnd,
// MinLoc/MaxLoc and the fused Sum+Min+Max+MinLoc of one forall).
//
// Reference: the HIP reducers fold per-thread values with a shuffle tree,
// per-block slots and a last-block-done pass, replicated 32 ways and folded
// again on the host (include/RAJA/policy/hip/reduce.hpp:127-208, 960-1161).
// Here kernel 1 computes the value of every fixed tile of the span and
// kernel 2 folds the tile values with the canonical 64-ary tree and applies
// the caller's init (reduce_tree.hpp).  The combine tree depends only on n
// -- not on the grid, the timing or whatever else shares the GPU -- and no
// floating-point atomics are used.  Two kernels instead of a
// last-block-done pass: on MI355X a release fence at device scope writes
// back the whole L2 of the block's XCD, and a same-address device atomic per
// block serialises across the 8 XCDs; the kernel boundary gives the
// visibility for free.
#include "reduce_tree.hpp"
#include "workspace.hpp"

namespace
{
using namespace RAJA::cdna4::dev;

constexpr int kFinalWaves = 16;
constexpr int kFinalMaxGroups = 1024;  // spans up to 2^26 tiles

// Whether the first tree level runs as its own kernel (rt::group_kernel):
// for accumulators wider than 8 bytes once the final block would fold more
// than one group per wave (the fused Sum/Min/Max/MinLoc over 2^28 FP64:
// 8192 tile values, 20.5 us in one block against 10.5 us with the first
// level spread over 128 blocks, launches included; same bits).
template<typename Op>
bool split_first_level(int64_t ntiles)
{
  return sizeof(typename Op::acc_t) > 8 && rc::ceil_div(ntiles, (int64_t)64) > kFinalWaves;
}

// Tile values the partials buffer holds: ntiles, plus the group values
// when the first level is split.
template<typename Op>
int64_t partial_slots(int64_t ntiles)
{
  return ntiles + (split_first_level<Op>(ntiles) ? rc::ceil_div(ntiles, (int64_t)64) : 0);
}

// The kernels over caller-provided partials (partial_slots values).
template<typename T, typename Op>
int launch_reduce_core(const T* x, int64_t n, typename Op::acc_t init, typename Op::acc_t* out,
                       typename Op::acc_t* partials, hipStream_t s, const char* name)
{
  const int64_t ntiles = rc::ceil_div(n, rt::tile_elems<T>());
  const int64_t max_blocks = (int64_t)rc::num_cus() * 8;
  rc::TimedScope ts(name, s);
  if (ntiles > 0)
  {
    const unsigned grid = (unsigned)(ntiles < max_blocks ? ntiles : max_blocks);
    hipLaunchKernelGGL((rt::tiles_kernel<T, Op>), dim3(grid), dim3(rt::kBlock), 0, s, x, n, ntiles,
                       partials);
    RC_HIP(hipGetLastError());
  }
  const typename Op::acc_t* vals = partials;
  int64_t count = ntiles;
  if (split_first_level<Op>(ntiles))
  {
    const int64_t groups = rc::ceil_div(ntiles, (int64_t)64);
    hipLaunchKernelGGL((rt::group_kernel<Op>), dim3((unsigned)groups), dim3(64), 0, s, partials,
                       ntiles, partials + ntiles);
    RC_HIP(hipGetLastError());
    vals = partials + ntiles;
    count = groups;
  }
  hipLaunchKernelGGL((rt::final_kernel<Op, kFinalWaves, kFinalMaxGroups>), dim3(1),
                     dim3(kFinalWaves * 64), 0, s, vals, (const uint32_t*)nullptr,
                     (const uint32_t*)nullptr, count, init, out, (const int64_t*)nullptr);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
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
  const int64_t ntiles = rc::ceil_div(n, rt::tile_elems<T>());
  if (ntiles > (int64_t)kFinalMaxGroups * 64)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "span too long for one reduction");
  }
  RC_TRY(rc::ensure_tmp(c, (size_t)(ntiles > 0 ? partial_slots<Op>(ntiles) : 1) * sizeof(acc_t)));
  return launch_reduce_core<T, Op>(x, n, init, out, (acc_t*)c->tmp, s, name);
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

namespace rws
{
size_t sum_f64_ws_bytes(int64_t n)
{
  return rc::align_up((size_t)(rc::ceil_div(n, rt::tile_elems<double>()) + 1) * sizeof(double), 256);
}

int sum_f64(const double* x, int64_t n, double* out, void* ws, hipStream_t s)
{
  if (rc::ceil_div(n, rt::tile_elems<double>()) > (int64_t)kFinalMaxGroups * 64)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "span too long for one reduction");
  }
  return launch_reduce_core<double, rk::OpPlus<double>>(x, n, 0.0, out, (double*)ws, s, "reduce");
}
}  // namespace rws

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
