// The canonical reduction tree shared by the span reductions and the
// segmented (batched) reductions.
//
// A range of n elements is cut into tiles of TILE elements from its start.
//   tile value : 256 threads; thread t folds vectors t, t+256, ..., (64 of
//                them, 16-byte each; vector k into accumulator k % 4, in
//                increasing k) into 4 interleaved accumulators, the
//                accumulators pairwise, then the block tree (DPP inside each
//                wave, the 4 wave values in order);
//   range value: a 64-ary tree over the tile values -- each level folds
//                groups of 64 consecutive values (identity-padded) with the
//                DPP wave tree -- then op(init, value).
// The tree is a function of n alone: any grid, any number of waves working
// on the tile values, and any batch the range is part of give the same
// bits.  A range reduced by the span entry point and the same range as one
// segment of a batch agree exactly.
#pragma once

#include "ops.hpp"
#include "runtime.hpp"

namespace rt
{
using namespace RAJA::cdna4::dev;

constexpr int kBlock = 256;
constexpr int kVecPerTile = 64;  // 16-byte vectors per thread per tile

// 16-byte loads a thread keeps in flight: fewer for operators whose
// accumulators are wide (8 for MinLoc/MaxLoc, 4 for the fused
// Sum/Min/Max/MinLoc), so the kernel keeps enough waves resident.  Element
// -> accumulator assignment (vector k goes to accumulator k % 4, in
// increasing k) does not depend on it, so neither do the result bits.
template<typename Op>
struct LoadsInFlight
{
  static constexpr int value =
      sizeof(typename Op::acc_t) > 24 ? 4 : (sizeof(typename Op::acc_t) > 8 ? 8 : 16);
};

template<typename T>
struct VecOf
{
  static constexpr int n = 16 / sizeof(T);
  typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
  // The same 16 bytes at the element's own alignment (a segment starts at
  // any element); 16-byte global accesses need only 4-byte alignment on CDNA.
  typedef T utype __attribute__((ext_vector_type(16 / sizeof(T)), aligned(sizeof(T))));
};

template<typename T>
constexpr int64_t tile_elems()
{
  return (int64_t)kBlock * kVecPerTile * VecOf<T>::n;
}

// acc <- op(acc, lift(x_i, i)); operators whose accumulators are wide give a
// cheaper accumulate() that is exact for elements met in increasing index
// order, which is how each accumulator meets them.
template<typename Op, typename T>
__device__ __forceinline__ auto accumulate_one(typename Op::acc_t& a, T v, int64_t i, int)
    -> decltype(Op::accumulate(a, v, i), void())
{
  Op::accumulate(a, v, i);
}
template<typename Op, typename T>
__device__ __forceinline__ void accumulate_one(typename Op::acc_t& a, T v, int64_t i, long)
{
  a = Op{}(a, Op::lift(v, i));
}

// Value of tile `tile` of the range x[0..n), in every thread of the block.
// The element -> (thread, accumulator) assignment does not depend on the
// alignment of x: loads are 16 bytes at the element's alignment.
template<typename T, typename Op>
__device__ __forceinline__ typename Op::acc_t fold_tile(const T* __restrict__ x, int64_t n,
                                                        int64_t tile,
                                                        typename Op::acc_t* s_scratch)
{
  using acc_t = typename Op::acc_t;
  using vec_t = typename VecOf<T>::type;
  using uvec_t = typename VecOf<T>::utype;
  constexpr int VEC = VecOf<T>::n;
  constexpr int64_t TILE = tile_elems<T>();
  constexpr int kVecPerChunk = LoadsInFlight<Op>::value;
  constexpr int kChunks = kVecPerTile / kVecPerChunk;
  static_assert(kVecPerChunk % 4 == 0 && kVecPerTile % kVecPerChunk == 0, "chunking keeps k % 4");
  const Op op{};
  const int t = (int)threadIdx.x;
  const int64_t base = tile * TILE;
  acc_t acc[4] = {Op::ident(), Op::ident(), Op::ident(), Op::ident()};
  if (base + TILE <= n)
  {
#pragma unroll 1
    for (int c = 0; c < kChunks; ++c)
    {
      vec_t v[kVecPerChunk];
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int k = c * kVecPerChunk + j;
        // read once: nontemporal (2^28 FP64 sum 0.363 -> 0.316 ms)
        v[j] = __builtin_nontemporal_load(reinterpret_cast<const uvec_t*>(x + base + ((int64_t)k * kBlock + t) * VEC));
      }
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int k = c * kVecPerChunk + j;
        const int64_t i0 = base + ((int64_t)k * kBlock + t) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
          accumulate_one<Op>(acc[j & 3], v[j][e], i0 + e, 0);
        }
      }
    }
  }
  else
  {
    // Partial tile: the same chunks, a vector load where the whole vector is
    // in range and element loads guarded by i < n at the end (all of a
    // chunk's loads in flight), same element order; the loop ends at the
    // first chunk that starts past n.
#pragma unroll 1
    for (int c = 0; c < kChunks; ++c)
    {
      T v[kVecPerChunk][VEC];
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int64_t i0 = base + ((int64_t)(c * kVecPerChunk + j) * kBlock + t) * VEC;
        if (i0 + VEC <= n)
        {
          const uvec_t q = *reinterpret_cast<const uvec_t*>(x + i0);
#pragma unroll
          for (int e = 0; e < VEC; ++e) v[j][e] = q[e];
        }
        else
        {
#pragma unroll
          for (int e = 0; e < VEC; ++e) v[j][e] = (i0 + e < n) ? x[i0 + e] : T(0);
        }
      }
#pragma unroll
      for (int j = 0; j < kVecPerChunk; ++j)
      {
        const int64_t i0 = base + ((int64_t)(c * kVecPerChunk + j) * kBlock + t) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
          if (i0 + e < n) accumulate_one<Op>(acc[j & 3], v[j][e], i0 + e, 0);
        }
      }
      if (base + (int64_t)(c + 1) * kVecPerChunk * kBlock * VEC >= n) break;
    }
  }
  acc_t a = op(op(acc[0], acc[1]), op(acc[2], acc[3]));
  return block_reduce(a, op, s_scratch);
}

// 64-ary tree over `count` tile values, evaluated by the block's waves
// (blockDim.x = 64 * waves).  s_a holds >= ceil(count/64) values, s_b >=
// ceil(count/4096) values.  Returns the range value in every thread.
template<typename Op>
__device__ __forceinline__ typename Op::acc_t tree_fold(const typename Op::acc_t* vals,
                                                        int64_t count,
                                                        typename Op::acc_t* s_a,
                                                        typename Op::acc_t* s_b)
{
  using acc_t = typename Op::acc_t;
  const Op op{};
  const int lane = lane_id();
  const int wave = wave_id();
  const int nw = (int)(blockDim.x / wave_size);
  if (count <= 0) return Op::ident();
  const acc_t* in = vals;
  acc_t* outs[2] = {s_a, s_b};
  int level = 0;
  while (count > 1)
  {
    const int64_t groups = (count + wave_size - 1) / wave_size;
    acc_t* out = outs[level & 1];
    for (int64_t g = wave; g < groups; g += nw)
    {
      const int64_t i = g * wave_size + lane;
      const acc_t v = (i < count) ? in[i] : Op::ident();
      const acc_t r = wave_reduce(v, op);
      if (lane == 0) out[g] = r;
    }
    __syncthreads();
    in = out;
    count = groups;
    ++level;
  }
  const acc_t r = in[0];
  __syncthreads();
  return r;
}

// Kernel 1 of a span reduction: tile values of x[0..n), grid-stride.
template<typename T, typename Op>
__global__ __launch_bounds__(kBlock) void tiles_kernel(const T* __restrict__ x, int64_t n,
                                                       int64_t ntiles,
                                                       typename Op::acc_t* __restrict__ partials)
{
  using acc_t = typename Op::acc_t;
  __shared__ acc_t s_scratch[max_waves];
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
  {
    const acc_t v = fold_tile<T, Op>(x, n, tile, s_scratch);
    if (threadIdx.x == 0) partials[tile] = v;
  }
}

// First level of the 64-ary tree on its own (one 64-lane block per group of
// 64 tile values): group g's value is the wave fold of values 64g..64g+63,
// exactly as tree_fold's first level computes it, so folding the group
// values with tree_fold gives the bits of folding the tile values.  Used
// for wide accumulators, whose wave fold is too slow to run group after
// group in one block.
template<typename Op>
__global__ __launch_bounds__(64) void group_kernel(const typename Op::acc_t* __restrict__ in,
                                                   int64_t count,
                                                   typename Op::acc_t* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * wave_size + threadIdx.x;
  const typename Op::acc_t v = (i < count) ? in[i] : Op::ident();
  const typename Op::acc_t r = wave_reduce(v, Op{});
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// Kernel 2: the 64-ary tree over the tile values of each segment (one
// block per segment; a span reduction is one segment) and the init;
// segment s's result goes to out[out_index[s]] (out[s] without an index).
template<typename Op, int WAVES, int MAX_GROUPS>
__global__ __launch_bounds__(WAVES * 64) void final_kernel(
    const typename Op::acc_t* __restrict__ partials, const uint32_t* __restrict__ seg_first,
    const uint32_t* __restrict__ seg_count, int64_t single_count, typename Op::acc_t init,
    typename Op::acc_t* __restrict__ out, const int64_t* __restrict__ out_index)
{
  using acc_t = typename Op::acc_t;
  __shared__ acc_t s_a[MAX_GROUPS];
  __shared__ acc_t s_b[(MAX_GROUPS + 63) / 64 + 1];
  const int64_t seg = blockIdx.x;
  const int64_t first = seg_first ? seg_first[seg] : 0;
  const int64_t count = seg_count ? seg_count[seg] : single_count;
  const acc_t r = tree_fold<Op>(partials + first, count, s_a, s_b);
  if (threadIdx.x == 0) out[out_index ? out_index[seg] : seg] = Op::with_init(init, r);
}

}  // namespace rt
