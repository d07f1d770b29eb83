// Device-side building blocks of the cdna4 policy family (gfx950, wave64).
//
// Counterpart of include/RAJA/policy/hip/intrinsics.hpp:44-431 in the
// reference (warp/block reductions, device-scope accessors), written for
// CDNA4 from scratch:
//   * cross-lane data movement uses DPP row operations and v_readlane, not
//     ds_bpermute shuffles (DPP moves stay in the VALU; 64-bit values move as
//     two 32-bit halves);
//   * every reduction tree is fixed (a butterfly inside 16-lane rows, then the
//     four row results in a fixed order, then waves in index order), so a
//     result depends only on the data and the launch shape, never on timing;
//   * cross-block publication uses agent-scope atomics, which the compiler
//     lowers with the cache maintenance needed between the 8 XCDs' L2s.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace RAJA
{
namespace cdna4
{
namespace dev
{

constexpr int wave_size = 64;
constexpr int max_block_size = 1024;
constexpr int max_waves = max_block_size / wave_size;

// DPP control words (GFX9 encoding).
enum : int
{
  dpp_quad_1032 = 0xB1,  // quad_perm [1,0,3,2]: lane ^ 1
  dpp_quad_2301 = 0x4E,  // quad_perm [2,3,0,1]: lane ^ 2
  dpp_row_shr1 = 0x111,
  dpp_row_shr2 = 0x112,
  dpp_row_shr4 = 0x114,
  dpp_row_shr8 = 0x118,
  dpp_row_mirror = 0x140,       // lane -> 15 - lane within a 16-lane row
  dpp_row_half_mirror = 0x141,  // lane -> 7 - lane within an 8-lane half row
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / wave_size); }

__device__ __forceinline__ uint64_t lanemask_lt()
{
  return (lane_id() == 0) ? 0ull : (~0ull >> (wave_size - lane_id()));
}

__device__ __forceinline__ uint64_t ballot(bool p)
{
  return __builtin_amdgcn_ballot_w64(p);
}

// DPP move of an arbitrary 4-byte-multiple value; lanes whose source lane is
// out of range keep `old`.
template<int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, typename T>
__device__ __forceinline__ T dpp_move(T old, T v)
{
  static_assert(sizeof(T) % 4 == 0, "dpp_move needs a 4-byte multiple");
  constexpr int N = sizeof(T) / 4;
  int o[N], s[N], r[N];
  __builtin_memcpy(o, &old, sizeof(T));
  __builtin_memcpy(s, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < N; ++k)
  {
    r[k] = __builtin_amdgcn_update_dpp(o[k], s[k], CTRL, ROW_MASK, BANK_MASK,
                                       false);
  }
  T res;
  __builtin_memcpy(&res, r, sizeof(T));
  return res;
}

// Value of `v` held by lane `src` (src must be wave-uniform).
template<typename T>
__device__ __forceinline__ T read_lane(T v, int src)
{
  static_assert(sizeof(T) % 4 == 0, "read_lane needs a 4-byte multiple");
  constexpr int N = sizeof(T) / 4;
  int s[N], r[N];
  __builtin_memcpy(s, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < N; ++k)
  {
    r[k] = __builtin_amdgcn_readlane(s[k], src);
  }
  T res;
  __builtin_memcpy(&res, r, sizeof(T));
  return res;
}

// Wave-wide reduction with a fixed tree; every lane returns the result.
// Requires all 64 lanes active.  `op` must be commutative (plus, min, max,
// bitwise, value-index pairs ordered lexicographically).
template<typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op)
{
  v = op(v, dpp_move<dpp_quad_1032>(v, v));
  v = op(v, dpp_move<dpp_quad_2301>(v, v));
  v = op(v, dpp_move<dpp_row_half_mirror>(v, v));
  v = op(v, dpp_move<dpp_row_mirror>(v, v));
  T r0 = read_lane(v, 0), r1 = read_lane(v, 16);
  T r2 = read_lane(v, 32), r3 = read_lane(v, 48);
  return op(op(r0, r1), op(r2, r3));
}

// Wave-wide scan.  Returns the inclusive prefix; `excl` receives the
// exclusive prefix (identity in lane 0) and `total` the wave total.
template<typename T, typename Op>
__device__ __forceinline__ T wave_scan(T v, Op op, T ident, T& excl, T& total)
{
  // Hillis-Steele inside each 16-lane row; lanes whose source falls outside
  // the row take the identity.
  v = op(dpp_move<dpp_row_shr1>(ident, v), v);
  v = op(dpp_move<dpp_row_shr2>(ident, v), v);
  v = op(dpp_move<dpp_row_shr4>(ident, v), v);
  v = op(dpp_move<dpp_row_shr8>(ident, v), v);
  // Row totals, combined in row order.
  const T r0 = read_lane(v, 15);
  const T r1 = read_lane(v, 31);
  const T r2 = read_lane(v, 47);
  const T r3 = read_lane(v, 63);
  const T c1 = r0;
  const T c2 = op(c1, r1);
  const T c3 = op(c2, r2);
  total = op(c3, r3);
  const int row = lane_id() >> 4;
  const T carry = (row == 0) ? ident : (row == 1) ? c1 : (row == 2) ? c2 : c3;
  // Exclusive: row-shifted inclusive value; the first lane of a row takes the
  // carry of the rows before it.
  const T row_excl = dpp_move<dpp_row_shr1>(ident, v);
  const bool row_first = (lane_id() & 15) == 0;
  excl = row_first ? carry : op(carry, row_excl);
  return (row == 0) ? v : op(carry, v);
}

// Block-wide reduction (blockDim.x a multiple of 64).  Every thread returns
// the block result.  `scratch` holds max_waves values in LDS.
template<typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T* scratch)
{
  v = wave_reduce(v, op);
  const int nw = (int)(blockDim.x / wave_size);
  if (lane_id() == 0)
  {
    scratch[wave_id()] = v;
  }
  __syncthreads();
  T r = scratch[0];
  for (int w = 1; w < nw; ++w)
  {
    r = op(r, scratch[w]);
  }
  __syncthreads();
  return r;
}

// ---- agent-scope (whole GPU, all XCDs) atomics ----------------------------

template<typename T>
__device__ __forceinline__ T atomic_load_agent(const T* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ T atomic_load_acquire_agent(const T* p)
{
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ void atomic_store_agent(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ void atomic_store_release_agent(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ T atomic_add_agent(T* p, T v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ T atomic_add_acq_rel_agent(T* p, T v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL,
                                __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded waits.  A spin on another block's publication ends after
// kSpinLimitTicks of the 100 MHz wall clock (about 2.7 s, far beyond any
// legitimate wait, which lasts at most one predecessor block's run time);
// the caller then records the failure in a device word the host can query
// and proceeds, so a broken publication protocol gives a reported error
// instead of a hung GPU.  The clock is read once per 256 polls.
constexpr uint64_t kSpinLimitTicks = 1ull << 28;

struct SpinBound
{
  uint64_t t0 = 0;
  uint32_t polls = 0;
  __device__ __forceinline__ bool expired()
  {
    if ((++polls & 255u) != 0u) return false;
    const uint64_t t = wall_clock64();
    if (t0 == 0) t0 = t;
    return t - t0 > kSpinLimitTicks;
  }
};

// Per-stream control block for single-pass kernels.  All counters return to
// zero when a kernel's last block finishes, so consecutive kernels on one
// stream (and replays of a captured graph) reuse it without any clearing.
struct LaunchControl
{
  uint32_t ticket;  // next tile id handed out
  uint32_t done;    // blocks finished
  uint32_t epoch;   // bumped by every scan kernel; tags lookback descriptors
  uint32_t pad[61];
};

// Draw a tile id in launch order (thread 0), broadcast through LDS.  Tile
// ids from a counter, not blockIdx, guarantee that every tile a block waits
// on belongs to a block that is already running.
__device__ __forceinline__ uint32_t draw_tile(LaunchControl* ctl, uint32_t* s_tile)
{
  if (threadIdx.x == 0)
  {
    *s_tile = atomic_add_agent(&ctl->ticket, 1u);
  }
  __syncthreads();
  return *s_tile;
}

// Mark this block finished; the last one resets the control block.  Must be
// called by thread 0 after the block's last global access.
__device__ __forceinline__ void finish_tiles(LaunchControl* ctl, uint32_t nblocks,
                                             bool bump_epoch)
{
  const uint32_t prev = atomic_add_acq_rel_agent(&ctl->done, 1u);
  if (prev == nblocks - 1)
  {
    atomic_store_agent(&ctl->ticket, 0u);
    atomic_store_agent(&ctl->done, 0u);
    if (bump_epoch)
    {
      atomic_store_agent(&ctl->epoch, atomic_load_agent(&ctl->epoch) + 1u);
    }
  }
}

}  // namespace dev
}  // namespace cdna4
}  // namespace RAJA
