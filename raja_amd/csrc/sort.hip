// Onesweep LSD radix sort of keys or (key, value) pairs.
//
// Reference: RAJA's HIP backend calls rocprim::radix_sort_keys/pairs over
// all key bits with double buffers and copies back when the result lands in
// the alternate buffer (include/RAJA/policy/hip/sort.hpp:77-495, 653-689);
// the semantics are RAJA's stable sort by operators::less / greater.
//
// Design (8-bit digits, P = 8*sizeof(key)/8 passes):
//   1. one read of all keys builds the P digit histograms (LDS counters per
//      block, wave-uniform digits added once per wave, then global atomics);
//   2. the host reads the histograms back for large inputs and skips every
//      pass whose digit is the same for all keys;
//   3. each pass is one kernel: each block owns one tile (16384 keys of 32
//      bits; 32 consecutive tiles of every 256 go to one XCD), loads its keys,
//      ranks them inside each wave by 8 ballots (peer lanes
//      with the same digit) against per-wave LDS digit counters, publishes
//      its per-digit counts, stages the keys in LDS in digit order, loads
//      its values, looks back over predecessor tiles for each digit (one
//      thread per digit, 8 predecessor rows in flight per step), then writes
//      the keys out so equal digits land in contiguous runs and stages and
//      writes the values the same way.  Each pass clears the
//      descriptors of the next one (two alternating arrays), so no clearing
//      launch or device counter is needed.
// Ranking follows input order inside a wave (item j, then lane) and across
// waves and tiles, so the sort is stable.
#include "ops.hpp"
#include "radix.hpp"
#include "runtime.hpp"
#include "segments.hpp"
#include "tune.hpp"
#include "workspace.hpp"

namespace
{
using namespace RAJA::cdna4::dev;

constexpr uint32_t kAgg = 1u, kInc = 2u;
constexpr uint32_t kCountMask = 0x3FFFFFFFu;
// A descriptor row: 256 digit words, padded to 264 words so rows of
// neighbouring tiles start in different cache-line sets.
constexpr int kRowStride = kRadix + 8;
// Predecessor rows a look-back step reads with all loads in flight.
constexpr int kLookbackRows = 8;
constexpr int64_t kMaxSortN = (int64_t)kCountMask;  // counts live in 30 bits
// Word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format,
// the other fields zero (stride 0: offsets and the size are in bytes).
constexpr int kBufferWord3 = 0x00020000;

// Lookback waits that outlived SpinBound's limit (see raja_cdna4_device_status).
static __device__ uint32_t g_spin_timeouts;

__global__ void clear_words_kernel(uint32_t* p, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
  {
    p[i] = 0u;
  }
}

// ---- 1. all-pass digit histograms ------------------------------------------
// Counts the N keys every lane of the wave holds (all of them valid) into
// s_hist, bin b's copy c at word b * C + c, lane l counting into copy l % C.
// A digit that is the same for all 64 * N keys is added once: the test is
// an OR over the keys of (bits ^ lane 0's first bits), one ballot per digit.
template<typename K, bool DESC, int P, int N, int C>
__device__ __forceinline__ void count_full_wave(const K (&k)[N], uint32_t* s_hist, int lane)
{
  using bits_t = typename KeyTraits<K>::bits_t;
  const bits_t b0 = KeyTraits<K>::to_bits(k[0]);
  bits_t f = 0;
  if constexpr (sizeof(bits_t) == 4)
  {
    f = (bits_t)__builtin_amdgcn_readfirstlane((int)b0);
  }
  else
  {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b0);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b0 >> 32));
    f = ((bits_t)hi << 32) | lo;
  }
  bits_t x = 0;
#pragma unroll
  for (int q = 0; q < N; ++q) x |= KeyTraits<K>::to_bits(k[q]) ^ f;
#pragma unroll
  for (int p = 0; p < P; ++p)
  {
    if (ballot(((x >> (8 * p)) & (bits_t)0xFF) != 0) == 0ull)
    {
      const uint32_t d0 = (uint32_t)((DESC ? ~f : f) >> (8 * p)) & (kRadix - 1);
      if (lane == 0) atomicAdd(&s_hist[(p * kRadix + d0) * C], (uint32_t)(wave_size * N));
    }
    else
    {
      uint32_t* const h = s_hist + p * kRadix * C + (lane % C);
#pragma unroll
      for (int q = 0; q < N; ++q) atomicAdd(&h[digit_of<K, DESC>(k[q], 8 * p) * C], 1u);
    }
  }
}

// Each wave streams 64 lanes x U x 16-byte vectors per iteration (U vector
// loads in flight per lane) and counts every digit of every key with LDS
// atomics; a digit shared by all keys of a full iteration is added once
// (constant high digits of small keys would otherwise serialise 64 lanes on
// one LDS word).
template<typename K, bool DESC, int P>
__global__ __launch_bounds__(256) void hist_kernel(const K* __restrict__ keys, int64_t n,
                                                   uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ desc0,
                                                   int64_t desc0_words)
{
  constexpr int VEC = 16 / sizeof(K);
  constexpr int U = 4;
  constexpr int PER_WAVE = wave_size * U * VEC;
  typedef K vec_t __attribute__((ext_vector_type(16 / sizeof(K))));
  // kCopies counters per bin, side by side (bin b's copy c is word
  // b * kCopies + c), lane l counting into copy l % kCopies: lanes of one
  // atomic instruction that meet the same bin use different words in
  // different banks.  2^28 keys: 0.283 ms with one copy, 0.239 ms with 4
  // copies as separate histograms, 0.229 / 0.212 / 0.27 ms with 4 / 8 / 16
  // side-by-side copies (copy l / 8 instead of l % 8: 0.285 ms).
  constexpr int kCopies = 8;
  __shared__ uint32_t s_hist[kCopies * P * kRadix];
  for (int i = threadIdx.x; i < kCopies * P * kRadix; i += 256) s_hist[i] = 0u;
  __syncthreads();
  const int lane = lane_id();
  const bool aligned = ((uintptr_t)keys % 16) == 0;
  const int64_t nwaves = (int64_t)gridDim.x * (256 / wave_size);
  for (int64_t base = ((int64_t)blockIdx.x * (256 / wave_size) + wave_id()) * PER_WAVE; base < n;
       base += nwaves * PER_WAVE)
  {
    K k[U * VEC];
    const bool full = aligned && base + PER_WAVE <= n;
    if (full)
    {
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        const vec_t x = __builtin_nontemporal_load(reinterpret_cast<const vec_t*>(keys + base + ((int64_t)u * wave_size + lane) * VEC));
#pragma unroll
        for (int e = 0; e < VEC; ++e) k[u * VEC + e] = x[e];
      }
    }
    else
    {
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
          const int64_t i = base + ((int64_t)u * wave_size + lane) * VEC + e;
          k[u * VEC + e] = (i < n) ? keys[i] : K(0);
        }
      }
    }
    if (full)
    {
      count_full_wave<K, DESC, P, U * VEC, kCopies>(k, s_hist, lane);
      continue;
    }
#pragma unroll
    for (int q = 0; q < U * VEC; ++q)
    {
      const int64_t i = base + ((int64_t)(q / VEC) * wave_size + lane) * VEC + (q % VEC);
      const bool valid = full || i < n;
      // valid lanes of one item form a prefix in lane order
      const uint64_t vmask = ballot(valid);
#pragma unroll
      for (int p = 0; p < P; ++p)
      {
        const uint32_t d = digit_of<K, DESC>(k[q], 8 * p);
        const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
        const bool uniform = (ballot(d == d0 || !valid) == ~0ull) && (vmask & 1ull);
        if (uniform)
        {
          if (lane == 0) atomicAdd(&s_hist[(p * kRadix + d0) * kCopies], (uint32_t)__popcll(vmask));
        }
        else if (valid)
        {
          atomicAdd(&s_hist[(p * kRadix + d) * kCopies], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * kRadix; i += 256)
  {
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kCopies; ++r) c += s_hist[i * kCopies + r];
    if (c) atomic_add_agent(&hist[i], c);
  }
  // Clear the descriptors the first pass will use.
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < desc0_words; i += stride)
  {
    desc0[i] = 0u;
  }
}

// ---- 2. one onesweep pass -----------------------------------------------------
// Workgroups are dispatched in increasing id order on each XCD and tiles are
// assigned in increasing order per XCD (see the tile mapping below), so every
// tile a block waits for is resident or done.  Keys, then
// values, go through one LDS staging buffer so a 16384-key tile (512
// threads x 32 items; 8192 keys, 16 items, when keys or values are 64-bit)
// fits in ~74 KiB, two blocks per CU, and the output is written in
// digit-contiguous runs.  The per-tile cost (barriers, digit
// scans, look-back) is paid once per 16384 keys: larger tiles measured
// faster than more resident blocks.  The keys are staged and the values
// loaded before the look-back, so both overlap its latency.
// SEG: the tiles of many ranges (one table record per tile, see
// rs::SortTile) sorted in one launch per pass.  keys_in / vals_in /
// keys_out / vals_out are then the bases of the input, output and
// alternate buffers in the order of `bases` below; a range reads the pass's
// source and writes its destination (pass 0: input -> alternate, odd
// passes: alternate -> output, even ones: output -> alternate), takes its
// digit offsets from its own histogram row and looks back only as far as
// its own first tile.
template<typename K, typename V, bool DESC, bool HAS_VALS, int BLOCK, int ITEMS, bool SEG = false>
__global__ __launch_bounds__(BLOCK) void onesweep_kernel(
    const K* __restrict__ keys_in, K* __restrict__ keys_out, const V* __restrict__ vals_in,
    V* __restrict__ vals_out, int64_t n, int shift, const uint32_t* __restrict__ hist,
    uint32_t* __restrict__ desc_cur, uint32_t* __restrict__ desc_next,
    const rs::SortTile* __restrict__ recs = nullptr, int pass = 0,
    K* __restrict__ keys_alt = nullptr, V* __restrict__ vals_alt = nullptr)
{
  static_assert(BLOCK >= kRadix && BLOCK % wave_size == 0, "one thread per digit");
  constexpr int NW = BLOCK / wave_size;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int WCHUNK = wave_size * ITEMS;
  constexpr int SLOT = (HAS_VALS && sizeof(V) > sizeof(K)) ? sizeof(V) : sizeof(K);
  __shared__ __attribute__((aligned(16))) unsigned char s_buf[TILE * SLOT];
  __shared__ uint32_t s_whist[NW * kRadix];
  __shared__ int32_t s_off[kRadix];
  __shared__ uint32_t s_tstart[kRadix];
  __shared__ uint32_t s_scan[2 * (kRadix / wave_size)];
  K* s_keys = reinterpret_cast<K*>(s_buf);
  V* s_vals = reinterpret_cast<V*>(s_buf);

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = wave_id();
  for (int i = tid; i < NW * kRadix; i += BLOCK) s_whist[i] = 0u;
  // Tile of this block.  Blocks go to the 8 XCDs round-robin (block b to
  // XCD b % 8), and each XCD has its own L2.  Mapping 32 consecutive tiles of
  // every group of 256 to one XCD lets that L2 merge the adjacent digit runs
  // of neighbouring tiles (tile t's run of digit d ends where tile t+1's
  // begins) before they reach HBM: one pass over 2^28 pairs took 1.276 ms
  // with tile = block, 0.997 ms with 8 per group of 64 and 0.985 ms with 32
  // per group of 256.  Each XCD dispatches its blocks in order and receives
  // its tiles in increasing order, so every block before the lowest
  // unfinished tile's block on that XCD is finished and that block is placed
  // as soon as the XCD has room (lookback launches run one at a time per
  // device, rc::LookbackGate); the last, partial group keeps tile = block.
  const int64_t tile = [] {
    constexpr int64_t KX = 32, GX = 8 * KX;
    const int64_t b = blockIdx.x, g = b / GX;
    if ((g + 1) * GX > (int64_t)gridDim.x) return b;
    const int64_t r = b - g * GX;
    return g * GX + (r & 7) * KX + (r >> 3);
  }();
  int64_t tj = tile;     // tile index inside its range
  int64_t tfirst = 0;    // descriptor row of the range's first tile
  if constexpr (SEG)
  {
    const rs::SortTile r = recs[tile];
    tj = r.j;
    tfirst = tile - r.j;
    n = r.len;
    hist += (int64_t)r.seg * (int64_t)sizeof(K) * kRadix;
    const K* ks = pass == 0 ? keys_in + r.ik : ((pass & 1) ? keys_alt + r.ak : keys_out + r.ok);
    K* kd = (pass & 1) ? keys_out + r.ok : keys_alt + r.ak;
    keys_in = ks;
    keys_out = kd;
    if (HAS_VALS)
    {
      const V* vs = pass == 0 ? vals_in + r.iv : ((pass & 1) ? vals_alt + r.av : vals_out + r.ov);
      V* vd = (pass & 1) ? vals_out + r.ov : vals_alt + r.av;
      vals_in = vs;
      vals_out = vd;
    }
  }
  const int64_t tbase = tj * TILE;
  const int64_t wbase = tbase + (int64_t)wave * WCHUNK;

  // Item j of this lane is element wbase + j*64 + lane; the wave holds
  // wvalid valid elements (32-bit compares below, no 64-bit indices kept).
  const int64_t wrem = n - wbase;
  const int wvalid = wrem <= 0 ? 0 : (wrem >= WCHUNK ? WCHUNK : (int)wrem);

  K k[ITEMS];
  const bool full = tbase + TILE <= n;
  if (full)
  {
    const K* kp = keys_in + wbase;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) k[j] = kp[j * wave_size + lane];
  }
  else
  {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      const int q = j * wave_size + lane;
      k[j] = (q < wvalid) ? keys_in[wbase + q] : K(0);
    }
  }
  for (int i = tid; i < kRowStride; i += BLOCK) desc_next[tile * kRowStride + i] = 0u;  // next pass
  __syncthreads();

  // Wave-local ranks: rank = digit count so far in this wave + lower peers.
  uint32_t rank[ITEMS];
  auto rank_items = [&](auto full_c) {
    constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      const bool valid = FULL || j * wave_size + lane < wvalid;
      const uint32_t d = digit_of<K, DESC>(k[j], shift);
      // Peers: the valid lanes whose digit equals this lane's, bit by bit.
      // sx is 0 or all ones (the bit sign-extended); lane m agrees with this
      // lane on the bit where ~(ballot ^ sx) has bit m set, so
      // p = p & ~(ballot ^ sx), one v_bitop3 per half, keeps the lanes that
      // agree on every bit.  A full tile starts from all lanes, a partial
      // one from the valid lanes.
      uint32_t pl = ~0u, ph = ~0u;
      if constexpr (!FULL)
      {
        const uint64_t vb = ballot(valid);
        pl = (uint32_t)vb;
        ph = (uint32_t)(vb >> 32);
      }
#pragma unroll
      for (int b = 0; b < 8; ++b)
      {
        int32_t sx = __builtin_amdgcn_sbfe((int32_t)d, b, 1);
        asm volatile("" : "+v"(sx));  // keep the one-instruction extract
        const uint64_t bal = ballot(sx != 0);
        pl = __builtin_amdgcn_bitop3_b32(pl, (uint32_t)bal, (uint32_t)sx, 0x90);
        ph = __builtin_amdgcn_bitop3_b32(ph, (uint32_t)(bal >> 32), (uint32_t)sx, 0x90);
      }
      // Peers in lower lanes: v_mbcnt counts the mask's bits below this lane.
      const uint32_t lower = __builtin_amdgcn_mbcnt_hi(ph, __builtin_amdgcn_mbcnt_lo(pl, 0u));
      uint32_t cnt = 0;
      if (valid) cnt = s_whist[wave * kRadix + d];
      rank[j] = (cnt + lower) | (d << 16);  // digit kept beside the rank
      asm volatile("" : "+v"(rank[j]));  // one register per item from here on
      if (valid && lower == 0)
      {
        s_whist[wave * kRadix + d] = cnt + (uint32_t)__builtin_popcount(pl) + (uint32_t)__builtin_popcount(ph);
      }
      __builtin_amdgcn_sched_barrier(0);  // rank one item at a time: few live masks
    }
  };
  if (full)
    rank_items(std::true_type{});
  else
    rank_items(std::false_type{});
  __syncthreads();

  // Per digit (thread = digit): wave offsets, tile count, publication and
  // digit scans.
  uint32_t ex_t = 0, ex_g = 0, prefix = 0, total = 0;
  uint32_t* const row = desc_cur + tile * kRowStride;
  if (tid < kRadix)
  {
    const int d = tid;
#pragma unroll
    for (int w = 0; w < NW; ++w)
    {
      const uint32_t c = s_whist[w * kRadix + d];
      s_whist[w * kRadix + d] = total;
      total += c;
    }
    atomic_store_agent(&row[d], ((tj == 0 ? kInc : kAgg) << 30) | total);
    const uint32_t gcount = hist[d];
    uint32_t tot_t, tot_g;
    (void)wave_scan(total, rk::OpPlus<uint32_t>{}, 0u, ex_t, tot_t);
    (void)wave_scan(gcount, rk::OpPlus<uint32_t>{}, 0u, ex_g, tot_g);
    if (lane == 0)
    {
      s_scan[wave] = tot_t;
      s_scan[kRadix / wave_size + wave] = tot_g;
    }
  }
  __syncthreads();
  uint32_t tstart = 0, gstart = 0;
  if (tid < kRadix)
  {
    tstart = ex_t;
    gstart = ex_g;
    for (int w = 0; w < wave; ++w)
    {
      tstart += s_scan[w];
      gstart += s_scan[kRadix / wave_size + w];
    }
    s_tstart[tid] = tstart;
  }
  __syncthreads();

  // Keys: stage in LDS in digit order (tile-local offsets only), before the
  // look-back, so the keys' registers are free while it runs.
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
  {
    if (j * wave_size + lane < wvalid)
    {
      const uint32_t dj = rank[j] >> 16;
      rank[j] = (rank[j] & 0xFFFFu) + s_tstart[dj] + s_whist[wave * kRadix + dj];  // tile position
      s_keys[rank[j]] = k[j];
    }
  }
  // Values are loaded now: their latency overlaps the look-back.
  V v[HAS_VALS ? ITEMS : 1];
  if (HAS_VALS)
  {
    if (full)
    {
      const V* vp = vals_in + wbase;
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) v[j] = vp[j * wave_size + lane];
    }
    else
    {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
      {
        const int q = j * wave_size + lane;
        v[j] = (q < wvalid) ? vals_in[wbase + q] : V(0);
      }
    }
  }
  if (tj > 0)
  {
    if (tid < kRadix)
    {
      // Descriptors tile-1, tile-2, ... are read kLookbackRows at a time
      // with all loads in flight; the fold stops at the first inclusive
      // count.  A descriptor not yet published is re-read.
      const int d = tid;
      bool done = false;
      SpinBound bound;
      for (int64_t t0 = tile - 1; !done; t0 -= kLookbackRows)
      {
        uint32_t w[kLookbackRows];
#pragma unroll
        for (int u = 0; u < kLookbackRows; ++u)
        {
          const int64_t t = t0 - u;
          w[u] = (t >= tfirst) ? atomic_load_agent(&desc_cur[t * kRowStride + d]) : (kInc << 30);
        }
#pragma unroll
        for (int u = 0; u < kLookbackRows; ++u)
        {
          if (done) break;
          const int64_t t = t0 - u;
          if (t < tfirst)
          {
            done = true;
            break;
          }
          while ((w[u] >> 30) == 0u)
          {
            __builtin_amdgcn_s_sleep(1);
            w[u] = atomic_load_agent(&desc_cur[t * kRowStride + d]);
            if ((w[u] >> 30) == 0u && bound.expired())
            {
              atomicAdd(&g_spin_timeouts, 1u);
              w[u] = kInc << 30;  // give up on it: the result is reported invalid
            }
          }
          prefix += w[u] & kCountMask;
          if ((w[u] >> 30) == kInc) done = true;
        }
      }
      atomic_store_agent(&row[d], (kInc << 30) | (prefix + total));
    }
  }
  if (tid < kRadix) s_off[tid] = (int32_t)(gstart + prefix) - (int32_t)tstart;
  __syncthreads();
  const int64_t rem = n - tbase;
  const int tile_n = (rem < TILE) ? (int)rem : TILE;
  // Digits of the tile positions this thread writes, four per register: the
  // value pass writes the same positions after the keys' LDS is reused.
  uint32_t dpk[HAS_VALS ? (ITEMS + 3) / 4 : 1];
#pragma unroll
  for (int q = 0; q < (HAS_VALS ? (ITEMS + 3) / 4 : 1); ++q) dpk[q] = 0u;
  // Groups of G positions: the LDS reads of a group are issued together
  // (positions past tile_n read stale LDS inside the buffer and are not
  // written), then the offsets, then the stores.  Full tiles skip the
  // position test.  Destinations are < n < 2^30 for consistent histograms;
  // an inconsistent one must not write out of range: 4-byte elements are
  // stored through a buffer resource of n elements, whose range check drops
  // a store at or past the end (no compare or branch per store), wider ones
  // test the index.
  constexpr int G = 8;
  static_assert(ITEMS % G == 0, "scatter groups");
  const uint32_t n32 = (uint32_t)n;
  const bool full_tile = tile_n == TILE;
  const __amdgpu_buffer_rsrc_t keys_rsrc =
      __builtin_amdgcn_make_buffer_rsrc(keys_out, (short)0, (int)(n32 * 4u), kBufferWord3);
  const __amdgpu_buffer_rsrc_t vals_rsrc =
      __builtin_amdgcn_make_buffer_rsrc(vals_out, (short)0, (int)(n32 * 4u), kBufferWord3);
  auto scatter_keys = [&](auto full_c) {
    constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
    for (int m0 = 0; m0 < ITEMS; m0 += G)
    {
      K key[G];
      uint32_t dg[G];
      int32_t off[G];
#pragma unroll
      for (int g = 0; g < G; ++g) key[g] = s_keys[tid + (m0 + g) * BLOCK];
#pragma unroll
      for (int g = 0; g < G; ++g)
      {
        dg[g] = digit_of<K, DESC>(key[g], shift);
        off[g] = s_off[dg[g]];
      }
#pragma unroll
      for (int g = 0; g < G; ++g)
      {
        const int m = m0 + g;
        const int i = tid + m * BLOCK;
        const uint32_t d = (uint32_t)(off[g] + i);
        if constexpr (sizeof(K) == 4)
        {
          if (FULL || i < tile_n)
          {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, key[g]), keys_rsrc, (int)(d << 2), 0, 0);
          }
        }
        else if ((FULL || i < tile_n) && d < n32)
        {
          keys_out[d] = key[g];
        }
        if (HAS_VALS) dpk[m >> 2] |= dg[g] << (8 * (m & 3));
      }
      __builtin_amdgcn_sched_barrier(0);  // one group in flight: bounded registers
    }
  };
  if (full_tile)
    scatter_keys(std::true_type{});
  else
    scatter_keys(std::false_type{});
  if (HAS_VALS)
  {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      if (j * wave_size + lane < wvalid) s_vals[rank[j]] = v[j];
    }
    __syncthreads();
    auto scatter_vals = [&](auto full_c) {
      constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
      for (int m0 = 0; m0 < ITEMS; m0 += G)
      {
        V val[G];
        int32_t off[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
        {
          const int m = m0 + g;
          val[g] = s_vals[tid + m * BLOCK];
          off[g] = s_off[(dpk[m >> 2] >> (8 * (m & 3))) & 0xFFu];
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
        {
          const int i = tid + (m0 + g) * BLOCK;
          const uint32_t d = (uint32_t)(off[g] + i);
          if constexpr (sizeof(V) == 4)
          {
            if (FULL || i < tile_n)
            {
              __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, val[g]), vals_rsrc, (int)(d << 2), 0, 0);
            }
          }
          else if ((FULL || i < tile_n) && d < n32)
          {
            vals_out[d] = val[g];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    if (full_tile)
      scatter_vals(std::true_type{});
    else
      scatter_vals(std::false_type{});
  }
}

template<typename K, typename V, bool HAS_VALS>
struct SortShape
{
  static constexpr bool wide = sizeof(K) > 4 || (HAS_VALS && sizeof(V) > 4);
  static constexpr int BLOCK = 512;
  static constexpr int ITEMS = wide ? 16 : 32;  // 8192 / 16384 keys per tile
  static constexpr int64_t TILE = (int64_t)BLOCK * ITEMS;
  // Keys per thread the tuner may choose instead (4-byte keys and values
  // only; 0 is the default): smaller tiles give mid-sized sorts enough
  // blocks to fill the chip.
  static constexpr int kChoices[3] = {0, wide ? 0 : 16, wide ? 0 : 8};
  static int64_t tile(int items) { return (int64_t)BLOCK * (items > 0 ? items : ITEMS); }
};

// Working memory of one sort: alternate key/value buffers, the digit
// histograms and two descriptor arrays.
template<typename K, typename V, bool HAS_VALS>
struct SortWs
{
  using S = SortShape<K, V, HAS_VALS>;
  static constexpr int P = (int)sizeof(K);
  int64_t n, ntiles, desc_words;
  size_t keys_bytes, vals_bytes, hist_bytes, desc_bytes;
  explicit SortWs(int64_t n_, int items = 0) : n(n_)
  {
    ntiles = rc::ceil_div(n, S::tile(items));
    desc_words = ntiles * kRowStride;
    keys_bytes = rc::align_up((size_t)n * sizeof(K), 256);
    vals_bytes = HAS_VALS ? rc::align_up((size_t)n * sizeof(V), 256) : 0;
    hist_bytes = rc::align_up((size_t)P * kRadix * sizeof(uint32_t), 256);
    desc_bytes = 2 * (size_t)desc_words * sizeof(uint32_t);
  }
  size_t tmp_bytes() const { return keys_bytes + vals_bytes + hist_bytes; }
  size_t total() const { return tmp_bytes() + desc_bytes; }
};

// ---- segmented onesweep: the tiles of many ranges, one launch per pass ----------
// Consecutive tile records one histogram block takes at most (a range's
// records are consecutive, so a block adds its counts to the global rows
// once per range it meets instead of once per tile), and the blocks a launch
// keeps at least before it gives a block more than one record: a block of
// 4 records takes about 30 us, so small tables take one record per block.
constexpr int kSegHistTiles = 4;
constexpr int64_t kSegHistMinBlocks = 2048;

// Digit histograms of every range of a tile table: block b takes records
// per_block * b .. + per_block - 1, counts their keys' digits in LDS
// and adds the counts to the range's P rows of 256 counters when the range
// changes or its records end; it clears each record's descriptor row of the
// first pass.  Keys are read as 16-byte vectors at the element's alignment
// (a tile starts at any element), 4 in flight per thread, and counted as in
// hist_kernel (8 side-by-side copies of each bin, a digit uniform over a
// wave's full iteration added once).  The bench's batched serving spent
// 14.2 ms in this kernel with one 4-byte load per thread per step and one
// histogram, 9.0 ms with 8 keys in flight and the side-by-side copies, and
// each of those blocks took one record.
template<typename K, bool DESC, int P>
__global__ __launch_bounds__(256) void seg_hist_kernel(const K* __restrict__ keys,
                                                       const rs::SortTile* __restrict__ recs,
                                                       int64_t ntiles, int per_block,
                                                       int64_t tile_elems,
                                                       uint32_t* __restrict__ hist,
                                                       uint32_t* __restrict__ desc0)
{
  static_assert(sizeof(K) == 4, "the segmented onesweep takes 4-byte keys");
  constexpr int kCopies = 8;
  constexpr int VEC = 4;
  constexpr int U = 4;
  constexpr int kBins = P * kRadix * kCopies;
  typedef K uvec_t __attribute__((ext_vector_type(VEC), aligned(sizeof(K))));
  __shared__ uint32_t s_hist[kBins];
  for (int i = threadIdx.x; i < kBins; i += 256) s_hist[i] = 0u;
  __syncthreads();
  const int lane = lane_id();
  const int64_t t0 = (int64_t)blockIdx.x * per_block;
  const int64_t t1 = (t0 + per_block < ntiles) ? t0 + per_block : ntiles;
  for (int64_t t = t0; t < t1; ++t)
  {
    const rs::SortTile r = recs[t];
    const int64_t first = (int64_t)r.j * tile_elems;
    const int64_t rem = r.len - first;
    const int cnt = (int)(rem < tile_elems ? rem : tile_elems);
    const K* const kp = keys + r.ik + first;
    const int nvec = cnt / VEC;
    // waves step through the tile's vectors in runs of U * 256
    for (int base = 0; base < nvec; base += U * 256)
    {
      K k[U * VEC];
      if (base + U * 256 <= nvec)
      {
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
          const uvec_t x = __builtin_nontemporal_load(
              reinterpret_cast<const uvec_t*>(kp + (int64_t)(base + u * 256 + (int)threadIdx.x) * VEC));
#pragma unroll
          for (int e = 0; e < VEC; ++e) k[u * VEC + e] = x[e];
        }
        count_full_wave<K, DESC, P, U * VEC, kCopies>(k, s_hist, lane);
      }
      else
      {
        uint32_t* const hc = s_hist + lane % kCopies;
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
          const int v = base + u * 256 + (int)threadIdx.x;
          if (v < nvec)
          {
            const uvec_t x = *reinterpret_cast<const uvec_t*>(kp + (int64_t)v * VEC);
#pragma unroll
            for (int e = 0; e < VEC; ++e)
            {
#pragma unroll
              for (int p = 0; p < P; ++p)
              {
                atomicAdd(&hc[(p * kRadix + digit_of<K, DESC>(x[e], 8 * p)) * kCopies], 1u);
              }
            }
          }
        }
      }
    }
    // the last cnt % VEC keys
    if ((int)threadIdx.x < cnt - nvec * VEC)
    {
      const K key = kp[(int64_t)nvec * VEC + threadIdx.x];
      uint32_t* const hc = s_hist + lane % kCopies;
#pragma unroll
      for (int p = 0; p < P; ++p) atomicAdd(&hc[(p * kRadix + digit_of<K, DESC>(key, 8 * p)) * kCopies], 1u);
    }
    for (int i = threadIdx.x; i < kRowStride; i += 256) desc0[t * kRowStride + i] = 0u;
    if (t + 1 == t1 || recs[t + 1].seg != r.seg)
    {
      // add the range's counts to its rows and clear them for the next range
      __syncthreads();
      uint32_t* const h = hist + (int64_t)r.seg * P * kRadix;
      for (int b = threadIdx.x; b < P * kRadix; b += 256)
      {
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < kCopies; ++q)
        {
          c += s_hist[b * kCopies + q];
          s_hist[b * kCopies + q] = 0u;
        }
        if (c) atomic_add_agent(&h[b], c);
      }
      __syncthreads();
    }
  }
}

// Sorts every range of a tile table (records built with tile SortShape::TILE)
// from the input into the output buffers, through the alternates; P is even,
// so the last pass lands in the output.  hist holds nranges * P * 256
// counters, desc 2 * ntiles descriptor rows; both need no clearing.
template<typename K, typename V, bool HAS_VALS, bool DESC>
int seg_sort_core(const K* in_k, const V* in_v, K* out_k, V* out_v, K* alt_k, V* alt_v,
                  const rs::SortTile* d_recs, int64_t ntiles, int64_t nranges, uint32_t* hist,
                  uint32_t* desc, hipStream_t s)
{
  using S = SortShape<K, V, HAS_VALS>;
  constexpr int P = (int)sizeof(K);
  static_assert(P % 2 == 0, "an even pass count ends in the output");
  if (ntiles <= 0) return RAJA_CDNA4_SUCCESS;
  if (ntiles > 0x7FFFFFFF) return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "too many tiles");
  rc::LookbackGate gate(s);  // one lookback launch at a time on the device
  RC_TRY(gate.status());
  rc::TimedScope ts("segmented_onesweep", s);
  const int64_t words = nranges * P * kRadix;
  const int64_t cb = rc::ceil_div(words, 256);
  hipLaunchKernelGGL(clear_words_kernel, dim3((unsigned)(cb < 1024 ? cb : 1024)), dim3(256), 0, s,
                     hist, words);
  const int per_block =
      (int)std::max<int64_t>(1, std::min<int64_t>(kSegHistTiles, ntiles / kSegHistMinBlocks));
  hipLaunchKernelGGL((seg_hist_kernel<K, DESC, P>), dim3((unsigned)rc::ceil_div(ntiles, (int64_t)per_block)),
                     dim3(256), 0, s, in_k, d_recs, ntiles, per_block, S::TILE, hist, desc);
  RC_HIP(hipGetLastError());
  uint32_t* const d0 = desc;
  uint32_t* const d1 = desc + ntiles * kRowStride;
  for (int p = 0; p < P; ++p)
  {
    hipLaunchKernelGGL((onesweep_kernel<K, V, DESC, HAS_VALS, S::BLOCK, S::ITEMS, true>),
                       dim3((unsigned)ntiles), dim3(S::BLOCK), 0, s, in_k, out_k, in_v, out_v,
                       (int64_t)0, 8 * p, hist + p * kRadix, (p & 1) ? d1 : d0, (p & 1) ? d0 : d1,
                       d_recs, p, alt_k, alt_v);
    RC_HIP(hipGetLastError());
  }
  return RAJA_CDNA4_SUCCESS;
}

// Sort (in_keys, in_vals) into (keys, vals); in_keys == keys sorts in place.
// tmp holds SortWs::tmp_bytes(), desc_mem SortWs::desc_bytes (its contents
// are ignored: the histogram kernel clears the first array and every pass
// clears the next).  With `pinned` (>= hist bytes of pinned host memory) the
// histograms are read back -- one stream synchronisation -- to skip passes
// whose digit is the same for every key.
template<typename K, typename V, bool HAS_VALS, bool DESC>
int sort_core(const K* in_keys, const V* in_vals, K* keys, V* vals, int64_t n, hipStream_t s,
              char* tmp, uint32_t* desc_mem, uint32_t* pinned, int items = 0)
{
  using S = SortShape<K, V, HAS_VALS>;
  constexpr int P = (int)sizeof(K);
  if (items != S::kChoices[1] && items != S::kChoices[2]) items = 0;
  const SortWs<K, V, HAS_VALS> w(n, items);
  K* alt_keys = (K*)tmp;
  V* alt_vals = HAS_VALS ? (V*)(tmp + w.keys_bytes) : nullptr;
  uint32_t* hist = (uint32_t*)(tmp + w.keys_bytes + w.vals_bytes);
  uint32_t* desc[2] = {desc_mem, desc_mem + w.desc_words};
  const bool inplace = (const void*)in_keys == (const void*)keys;

  {
    rc::TimedScope ts("radix_hist", s);
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(256), 0, s, hist, (int64_t)P * kRadix);
    const int64_t want = rc::ceil_div(n, (int64_t)256 * 64);
    const int64_t cap = (int64_t)rc::num_cus() * 16;
    const unsigned grid = (unsigned)(want < 1 ? 1 : (want < cap ? want : cap));
    hipLaunchKernelGGL((hist_kernel<K, DESC, P>), dim3(grid), dim3(256), 0, s, in_keys, n, hist,
                       desc[0], w.desc_words);
    RC_HIP(hipGetLastError());
  }
  bool active[P];
  int npass = P;
  for (int p = 0; p < P; ++p) active[p] = true;
  if (pinned && n >= (int64_t)1 << 18)
  {
    RC_HIP(hipMemcpyAsync(pinned, hist, (size_t)P * kRadix * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < P; ++p)
    {
      for (int dd = 0; dd < kRadix; ++dd)
      {
        if ((int64_t)pinned[p * kRadix + dd] == n)
        {
          active[p] = false;
          --npass;
        }
      }
    }
  }
  rc::LookbackGate gate(s);  // one lookback launch at a time on the device
  RC_TRY(gate.status());
  // Out of place, an even pass count starts into the alternate buffers and
  // an odd one into the outputs, so the last pass always lands in the
  // outputs.  In place, passes start into the alternates and an odd count
  // ends with one copy back.
  const K* ksrc = in_keys;
  const V* vsrc = in_vals;
  bool to_alt = inplace || (npass % 2 == 0);
  int q = 0;
  for (int p = 0; p < P; ++p)
  {
    if (!active[p]) continue;
    K* kdst = to_alt ? alt_keys : keys;
    V* vdst = to_alt ? alt_vals : vals;
    rc::TimedScope ts("onesweep_pass", s);
    if (items != 0 && items == S::kChoices[1])
    {
      hipLaunchKernelGGL((onesweep_kernel<K, V, DESC, HAS_VALS, S::BLOCK, S::kChoices[1] ? S::kChoices[1] : S::ITEMS>),
                         dim3((unsigned)w.ntiles), dim3(S::BLOCK), 0, s, (K*)ksrc, kdst, (V*)vsrc,
                         vdst, n, 8 * p, hist + p * kRadix, desc[q & 1], desc[(q + 1) & 1]);
    }
    else if (items != 0 && items == S::kChoices[2])
    {
      hipLaunchKernelGGL((onesweep_kernel<K, V, DESC, HAS_VALS, S::BLOCK, S::kChoices[2] ? S::kChoices[2] : S::ITEMS>),
                         dim3((unsigned)w.ntiles), dim3(S::BLOCK), 0, s, (K*)ksrc, kdst, (V*)vsrc,
                         vdst, n, 8 * p, hist + p * kRadix, desc[q & 1], desc[(q + 1) & 1]);
    }
    else
    {
      hipLaunchKernelGGL((onesweep_kernel<K, V, DESC, HAS_VALS, S::BLOCK, S::ITEMS>),
                         dim3((unsigned)w.ntiles), dim3(S::BLOCK), 0, s, (K*)ksrc, kdst, (V*)vsrc,
                         vdst, n, 8 * p, hist + p * kRadix, desc[q & 1], desc[(q + 1) & 1]);
    }
    RC_HIP(hipGetLastError());
    ksrc = kdst;
    vsrc = vdst;
    to_alt = !to_alt;
    ++q;
  }
  if (ksrc != keys)
  {
    RC_HIP(hipMemcpyAsync(keys, ksrc, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if (HAS_VALS)
    {
      RC_HIP(hipMemcpyAsync(vals, vsrc, (size_t)n * sizeof(V), hipMemcpyDeviceToDevice, s));
    }
  }
  return RAJA_CDNA4_SUCCESS;
}

// In-place sort with the stream's scratch, in the tuned tile shape of its
// (key type, length bucket) (items < 0: look it up; else this shape).
template<typename K, typename V, bool HAS_VALS, bool DESC>
int sort_impl(K* keys, V* vals, int64_t n, hipStream_t s, int items = -1)
{
  if (n > kMaxSortN)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "sort length %lld exceeds %lld",
                         (long long)n, (long long)kMaxSortN);
  }
  using S = SortShape<K, V, HAS_VALS>;
  if (items < 0)
  {
    items = 0;
    if (!S::wide) RC_TRY(rtune::choice(rtune::kSortPairs, rk::dtype_code<K>(), n, s, &items));
  }
  if (items != S::kChoices[1] && items != S::kChoices[2]) items = 0;
  rc::StreamCtx* c = nullptr;
  RC_TRY(rc::get_ctx(s, &c));
  const SortWs<K, V, HAS_VALS> w(n, items);
  RC_TRY(rc::ensure_tmp(c, w.tmp_bytes()));
  RC_TRY(rc::ensure_sort_desc(c, w.desc_bytes));
  // Under stream capture the histogram readback (a host wait) cannot run:
  // every pass is executed instead, with the same result.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RC_HIP(hipStreamIsCapturing(s, &cap));
  uint32_t* const pinned = cap == hipStreamCaptureStatusNone ? (uint32_t*)c->pinned : nullptr;
  return sort_core<K, V, HAS_VALS, DESC>(keys, vals, keys, vals, n, s, (char*)c->tmp,
                                         (uint32_t*)c->sort_desc, pinned, items);
}

template<typename K>
int sort_key_type(void* keys, void* vals, int val_bytes, int64_t n, int descending, hipStream_t s)
{
  if (!vals)
  {
    return descending ? sort_impl<K, uint32_t, false, true>((K*)keys, nullptr, n, s)
                      : sort_impl<K, uint32_t, false, false>((K*)keys, nullptr, n, s);
  }
  if (val_bytes == 4)
  {
    return descending ? sort_impl<K, uint32_t, true, true>((K*)keys, (uint32_t*)vals, n, s)
                      : sort_impl<K, uint32_t, true, false>((K*)keys, (uint32_t*)vals, n, s);
  }
  if (val_bytes == 8)
  {
    return descending ? sort_impl<K, uint64_t, true, true>((K*)keys, (uint64_t*)vals, n, s)
                      : sort_impl<K, uint64_t, true, false>((K*)keys, (uint64_t*)vals, n, s);
  }
  return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "value size %d (4 or 8 supported)", val_bytes);
}

}  // namespace

namespace rsort
{
void tile_choices(int key_dtype, int out[3])
{
  const int* c = (key_dtype == RAJA_CDNA4_U32 || key_dtype == RAJA_CDNA4_I32 ||
                  key_dtype == RAJA_CDNA4_F32)
                     ? SortShape<uint32_t, uint32_t, true>::kChoices
                     : SortShape<uint64_t, uint32_t, true>::kChoices;
  for (int i = 0; i < 3; ++i) out[i] = c[i];
}

int tune_sort(int key_dtype, void* keys, void* vals, int64_t n, int items, hipStream_t s)
{
  switch (key_dtype)
  {
    case RAJA_CDNA4_U32: return sort_impl<uint32_t, uint32_t, true, false>((uint32_t*)keys, (uint32_t*)vals, n, s, items);
    case RAJA_CDNA4_I32: return sort_impl<int32_t, uint32_t, true, false>((int32_t*)keys, (uint32_t*)vals, n, s, items);
    case RAJA_CDNA4_F32: return sort_impl<float, uint32_t, true, false>((float*)keys, (uint32_t*)vals, n, s, items);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "sort shapes are tuned for 4-byte keys");
  }
}

int64_t seg_tile() { return SortShape<uint32_t, uint32_t, true>::TILE; }

size_t seg_desc_bytes(int64_t ntiles)
{
  return 2 * (size_t)ntiles * kRowStride * sizeof(uint32_t);
}

size_t seg_hist_bytes(int64_t nranges)
{
  return (size_t)nranges * 4 * kRadix * sizeof(uint32_t);
}

int seg_sort_table(int key_dtype, int val_bytes, int descending, const void* in_k, const void* in_v,
                   void* out_k, void* out_v, void* alt_k, void* alt_v, const rs::SortTile* d_recs,
                   int64_t ntiles, int64_t nranges, uint32_t* hist, uint32_t* desc, hipStream_t s)
{
  if (val_bytes != 0 && val_bytes != 4)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "segmented onesweep takes 4-byte values");
  }
#define RSORT_SEG_CASE(CODE, K)                                                                   \
  case CODE:                                                                                      \
    if (val_bytes == 4)                                                                           \
      return descending                                                                           \
                 ? seg_sort_core<K, uint32_t, true, true>(                                        \
                       (const K*)in_k, (const uint32_t*)in_v, (K*)out_k, (uint32_t*)out_v,        \
                       (K*)alt_k, (uint32_t*)alt_v, d_recs, ntiles, nranges, hist, desc, s)        \
                 : seg_sort_core<K, uint32_t, true, false>(                                       \
                       (const K*)in_k, (const uint32_t*)in_v, (K*)out_k, (uint32_t*)out_v,        \
                       (K*)alt_k, (uint32_t*)alt_v, d_recs, ntiles, nranges, hist, desc, s);       \
    return descending                                                                             \
               ? seg_sort_core<K, uint32_t, false, true>((const K*)in_k, nullptr, (K*)out_k,      \
                                                          nullptr, (K*)alt_k, nullptr, d_recs,    \
                                                          ntiles, nranges, hist, desc, s)         \
               : seg_sort_core<K, uint32_t, false, false>((const K*)in_k, nullptr, (K*)out_k,     \
                                                           nullptr, (K*)alt_k, nullptr, d_recs,   \
                                                           ntiles, nranges, hist, desc, s);
  switch (key_dtype)
  {
    RSORT_SEG_CASE(RAJA_CDNA4_U32, uint32_t)
    RSORT_SEG_CASE(RAJA_CDNA4_I32, int32_t)
    RSORT_SEG_CASE(RAJA_CDNA4_F32, float)
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "segmented onesweep takes 4-byte keys");
  }
#undef RSORT_SEG_CASE
}

int take_spin_timeouts(uint32_t* count)
{
  uint32_t v = 0;
  RC_HIP(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_spin_timeouts), sizeof(v)));
  if (v)
  {
    const uint32_t zero = 0;
    RC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_spin_timeouts), &zero, sizeof(zero)));
  }
  *count = v;
  return RAJA_CDNA4_SUCCESS;
}
}  // namespace rsort

namespace rws
{
size_t sort_u32_i32_ws_bytes(int64_t n)
{
  return SortWs<uint32_t, uint32_t, true>(n).total();
}

int sort_pairs_u32_i32(const uint32_t* in_keys, const int32_t* in_vals, uint32_t* out_keys,
                       int32_t* out_vals, int64_t n, void* ws, hipStream_t s)
{
  if (n <= 0) return RAJA_CDNA4_SUCCESS;
  if (n > kMaxSortN) return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "sort too long");
  if (n == 1)
  {
    RC_HIP(hipMemcpyAsync(out_keys, in_keys, 4, hipMemcpyDeviceToDevice, s));
    RC_HIP(hipMemcpyAsync(out_vals, in_vals, 4, hipMemcpyDeviceToDevice, s));
    return RAJA_CDNA4_SUCCESS;
  }
  const SortWs<uint32_t, uint32_t, true> w(n);
  return sort_core<uint32_t, uint32_t, true, false>(in_keys, (const uint32_t*)in_vals, out_keys,
                                                    (uint32_t*)out_vals, n, s, (char*)ws,
                                                    (uint32_t*)((char*)ws + w.tmp_bytes()), nullptr);
}
}  // namespace rws

extern "C" {

int raja_cdna4_sort_pairs(int key_dtype, void* keys, void* vals, int val_bytes, int64_t n,
                          int descending, void* stream)
{
  if (n < 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "negative length");
  if (n <= 1) return RAJA_CDNA4_SUCCESS;  // pattern/sort.hpp:74-82
  if (!keys) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "keys is NULL");
  hipStream_t s = (hipStream_t)stream;
  switch (key_dtype)
  {
    case RAJA_CDNA4_U32: return sort_key_type<uint32_t>(keys, vals, val_bytes, n, descending, s);
    case RAJA_CDNA4_I32: return sort_key_type<int32_t>(keys, vals, val_bytes, n, descending, s);
    case RAJA_CDNA4_F32: return sort_key_type<float>(keys, vals, val_bytes, n, descending, s);
    case RAJA_CDNA4_U64: return sort_key_type<uint64_t>(keys, vals, val_bytes, n, descending, s);
    case RAJA_CDNA4_I64: return sort_key_type<int64_t>(keys, vals, val_bytes, n, descending, s);
    case RAJA_CDNA4_F64: return sort_key_type<double>(keys, vals, val_bytes, n, descending, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "key dtype %d", key_dtype);
  }
}

}  // extern "C"
