// Single-pass decoupled-lookback scans (inclusive / exclusive, plus / min /
// max; the in-place variants are in == out).
//
// Reference: RAJA's HIP backend delegates to rocprim::inclusive_scan /
// exclusive_scan (include/RAJA/policy/hip/scan.hpp:51-274); the semantics
// are those of include/RAJA/policy/sequential/scan.hpp:43-153.  This kernel
// reads every element once and writes it once (8 B/elem for int32):
//   * a tile is 8 waves x 64 lanes x V vectors of 16 bytes (32768 int32 or
//     12288 int64 for a span, 16384 or 8192 int32 in tile tables); wave w owns a
//     contiguous eighth, lane l loads vector k at ((k*64 + l) * VEC), so
//     every load and store is a fully coalesced dwordx4;
//   * each vector is scanned in registers, vector totals are scanned across
//     the wave with DPP (one wave scan per k), waves are combined in LDS;
//   * the tile aggregate is published at once, then wave 0 looks back over
//     up to 64 predecessor tiles per step (one lane per tile) until it meets
//     an inclusive prefix;
//   * tile t is block t (no ticket counter: a same-address device atomic
//     per block serialises across the 8 XCDs at ~11 ns each);
//   * descriptors carry a per-stream epoch chosen by the host, so they never
//     need clearing between calls; under graph capture (where the epoch is
//     frozen into the graph) a clear kernel runs first.
#include <mutex>
#include <type_traits>

#include "ops.hpp"
#include "runtime.hpp"
#include "segments.hpp"
#include "tune.hpp"
#include "workspace.hpp"

namespace
{
using namespace RAJA::cdna4::dev;

constexpr int kBlock = 512;
constexpr int kWaves = kBlock / wave_size;
constexpr uint32_t kAgg = 1u, kInc = 2u;
// Segment plans of 4-byte values whose mean segment length is below this
// take the short-segment tile shape.
constexpr int64_t kShortMeanLen = 65536;

// Tile shapes: MODE 0 a span, MODE 1 a tile table (batched requests, long
// segments), MODE 2 a tile table of mostly short segments.
// VO != 0 overrides the vectors per lane of a span (a tuned shape).
template<typename T, int MODE = 0, int VO = 0>
struct ScanShape
{
  static constexpr int VEC = 16 / sizeof(T);
  // vectors per lane: for a span, 16 for 4-byte values (32768 per tile) and
  // 12 for 8-byte values (12288 per tile); with nontemporal loads and stores
  // these measured best (2^28 int32: 0.385 ms at 16, 0.43 at 8, 0.40 at 20;
  // int64: 0.83 ms at 12).  Tile tables keep 8 for 4-byte values, and 4 when
  // their segments are short on average: most of their tiles are partial
  // (20000 log-uniform segments of mean length 5800, int32: 0.160 ms at 4,
  // 0.170 at 6, 0.185 at 8, 0.200 at 12, 0.22 at 16).
  static constexpr int V =
      VO != 0 ? VO : (sizeof(T) == 8 ? 12 : (MODE == 0 ? 16 : (MODE == 1 ? 8 : 4)));
  static constexpr int WCHUNK = wave_size * V * VEC;   // elements per wave
  static constexpr int64_t TILE = (int64_t)kWaves * WCHUNK;
  typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
  // The same 16 bytes at the element's own alignment: a tile of a segment
  // starts at any element.  16-byte global accesses need only 4-byte
  // alignment on CDNA.
  typedef T uvec_t __attribute__((ext_vector_type(16 / sizeof(T)), aligned(sizeof(T))));
};

// Elements [i0, i0 + VEC) of a lane's vector where only indices < n exist:
// one 16-byte access when the whole vector is in range, element accesses at
// the range's end.
template<typename S, typename T>
__device__ __forceinline__ void load_vec_guarded(const T* in, int64_t i0, int64_t n,
                                                 T (&v)[S::VEC], T ident)
{
  if (i0 + S::VEC <= n)
  {
    const typename S::uvec_t x = *reinterpret_cast<const typename S::uvec_t*>(in + i0);
#pragma unroll
    for (int e = 0; e < S::VEC; ++e) v[e] = x[e];
  }
  else
  {
#pragma unroll
    for (int e = 0; e < S::VEC; ++e) v[e] = (i0 + e < n) ? in[i0 + e] : ident;
  }
}

template<typename S, typename T>
__device__ __forceinline__ void store_vec_guarded(T* out, int64_t i0, int64_t n,
                                                  const T (&o)[S::VEC])
{
  if (i0 + S::VEC <= n)
  {
    typename S::uvec_t x;
#pragma unroll
    for (int e = 0; e < S::VEC; ++e) x[e] = o[e];
    *reinterpret_cast<typename S::uvec_t*>(out + i0) = x;
  }
  else
  {
#pragma unroll
    for (int e = 0; e < S::VEC; ++e)
    {
      if (i0 + e < n) out[i0 + e] = o[e];
    }
  }
}

// Descriptor of one tile: one 64-bit word per 32-bit half of the value, each
// holding (status << 32 | half).  status = epoch << 2 | kind.
template<typename T>
struct Desc
{
  static constexpr int W = sizeof(T) / 4;
  uint64_t w[W];
};

template<typename T>
__device__ __forceinline__ void publish(Desc<T>* d, uint32_t status, T v)
{
  uint32_t h[Desc<T>::W];
  __builtin_memcpy(h, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < Desc<T>::W; ++k)
  {
    atomic_store_agent(&d->w[k], ((uint64_t)status << 32) | h[k]);
  }
}

// Returns the kind (0 if not yet published in this epoch) and the value.
template<typename T>
__device__ __forceinline__ uint32_t peek(const Desc<T>* d, uint32_t epoch_tag, T& v)
{
  uint32_t h[Desc<T>::W];
  uint32_t st0 = 0;
  bool same = true;
#pragma unroll
  for (int k = 0; k < Desc<T>::W; ++k)
  {
    const uint64_t w = atomic_load_agent(&d->w[k]);
    const uint32_t st = (uint32_t)(w >> 32);
    if (k == 0) st0 = st;
    same = same && (st == st0);
    h[k] = (uint32_t)w;
  }
  __builtin_memcpy(&v, h, sizeof(T));
  if (!same || (st0 & ~3u) != epoch_tag) return 0u;
  return st0 & 3u;
}

// Lookback waits that outlived SpinBound's limit (see raja_cdna4_device_status).
static __device__ uint32_t g_spin_timeouts;

template<typename T, typename Op>
__device__ __forceinline__ T lookback(const Desc<T>* desc, int64_t tile, int64_t j,
                                      uint32_t epoch_tag)
{
  // Predecessor `back` tiles behind belongs to this range while back <= j
  // (j = the tile's index inside its range; j = tile for a single span).
  const Op op{};
  const int lane = lane_id();
  T acc = Op::ident();
  int64_t step = 1;
  SpinBound bound;
  while (true)
  {
    const int64_t back = step + lane;
    const int64_t t = tile - back;
    T v = Op::ident();
    uint32_t kind = kInc;
    if (back <= j)
    {
      kind = peek<T>(desc + t, epoch_tag, v);
      while (ballot(kind == 0u) != 0ull)
      {
        __builtin_amdgcn_s_sleep(1);
        if (kind == 0u) kind = peek<T>(desc + t, epoch_tag, v);
        if (bound.expired())  // the loop and the clock are wave-uniform
        {
          if (lane == 0) atomicAdd(&g_spin_timeouts, 1u);
          if (kind == 0u)
          {
            kind = kInc;  // give up on it: the result is reported invalid
            v = Op::ident();
          }
        }
      }
    }
    else
    {
      v = Op::ident();
    }
    const uint64_t inc = ballot(kind == kInc);
    if (inc)
    {
      const int first = __builtin_ctzll(inc);
      const T contrib = (lane <= first) ? v : Op::ident();
      acc = op(wave_reduce(contrib, op), acc);
      return acc;
    }
    acc = op(wave_reduce(v, op), acc);
    step += wave_size;
  }
}

// Tile t is block t.  Workgroups are dispatched in increasing id order on
// each XCD, so every tile a block waits for is already resident or done.
// MODE != 0 (SEG): the tiles of many ranges (segments) come from a tile
// table; each range starts from `init` and its lookback stops at its own
// first tile.
template<typename T, typename Op, bool EXCLUSIVE, bool ALIGNED, int MODE, int VO = 0>
__global__ __launch_bounds__(kBlock) void scan_kernel(const T* in, T* out, int64_t n, T init,
                                                      Desc<T>* __restrict__ desc,
                                                      uint32_t epoch_tag,
                                                      const rs::TileRec* __restrict__ tiles)
{
  constexpr bool SEG = MODE != 0;
  using S = ScanShape<T, MODE, VO>;
  using vec_t = typename S::vec_t;
  constexpr int VEC = S::VEC;
  constexpr int V = S::V;
  const Op op{};
  __shared__ T s_wave_agg[kWaves];
  __shared__ T s_prefix;

  const int64_t tile = blockIdx.x;
  int64_t j = tile;               // tile index inside its range
  int64_t tbase = tile * S::TILE;  // first element of the tile
  int64_t nend = n;               // end of the range
  if (SEG)
  {
    const rs::TileRec r = tiles[tile];
    j = r.j;
    tbase = r.seg_start + (int64_t)r.j * S::TILE;
    nend = r.seg_start + r.seg_len;
    out += r.out_delta;  // output element i of the range lands at out[i + out_delta]
  }
  const int lane = lane_id();
  const int wave = wave_id();
  const int64_t wbase = tbase + (int64_t)wave * S::WCHUNK;
  // A segment's tiles start at any element: they take 16-byte accesses at the
  // element's alignment.  A span's tiles take 16-byte aligned ones when the
  // host found both pointers aligned.
  using lvec_t = std::conditional_t<SEG, typename S::uvec_t, vec_t>;
  const bool full = (SEG || ALIGNED) && tbase + S::TILE <= nend;
  n = nend;

  T v[V][VEC];
  if (full)
  {
#pragma unroll
    for (int k = 0; k < V; ++k)
    {
      // streamed once in and once out: nontemporal loads and stores (about 3.5 % faster)
      const lvec_t x = __builtin_nontemporal_load(reinterpret_cast<const lvec_t*>(in + wbase + ((int64_t)k * wave_size + lane) * VEC));
#pragma unroll
      for (int e = 0; e < VEC; ++e) v[k][e] = x[e];
    }
  }
  else
  {
#pragma unroll
    for (int k = 0; k < V; ++k)
    {
      load_vec_guarded<S>(in, wbase + ((int64_t)k * wave_size + lane) * VEC, n, v[k], Op::ident());
    }
  }

  // Per-lane exclusive prefix of each vector within the wave's chunk.
  T lane_pref[V];
  T run = Op::ident();
#pragma unroll
  for (int k = 0; k < V; ++k)
  {
    T t = v[k][0];
#pragma unroll
    for (int e = 1; e < VEC; ++e) t = op(t, v[k][e]);
    T excl, total;
    (void)wave_scan(t, op, Op::ident(), excl, total);
    lane_pref[k] = op(run, excl);
    run = op(run, total);
  }
  if (lane == 0) s_wave_agg[wave] = run;
  __syncthreads();
  T wave_pref = Op::ident();
  T tile_agg = Op::ident();
#pragma unroll
  for (int w = 0; w < kWaves; ++w)
  {
    if (w < wave) wave_pref = op(wave_pref, s_wave_agg[w]);
    tile_agg = op(tile_agg, s_wave_agg[w]);
  }
  const T start = EXCLUSIVE ? init : Op::ident();
  if (wave == 0)
  {
    T prefix;
    if (j == 0)
    {
      prefix = start;
      if (lane == 0) publish<T>(desc + tile, epoch_tag | kInc, op(prefix, tile_agg));
    }
    else
    {
      if (lane == 0) publish<T>(desc + tile, epoch_tag | kAgg, tile_agg);
      // 8-byte span scans: the lookback wave at raised issue priority (2^28
      // int64 0.821 -> 0.805 ms; no gain for 4-byte or segmented scans).
      constexpr bool kPrio = sizeof(T) == 8 && MODE == 0;
      if (kPrio) __builtin_amdgcn_s_setprio(3);
      // Inclusive descriptors already hold `start` (tile 0 folds it in).
      prefix = lookback<T, Op>(desc, tile, j, epoch_tag);
      if (lane == 0) publish<T>(desc + tile, epoch_tag | kInc, op(prefix, tile_agg));
      if (kPrio) __builtin_amdgcn_s_setprio(0);
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const T base_pref = op(s_prefix, wave_pref);

  T o[V][VEC];
#pragma unroll
  for (int k = 0; k < V; ++k)
  {
    T acc = op(base_pref, lane_pref[k]);
#pragma unroll
    for (int e = 0; e < VEC; ++e)
    {
      if (EXCLUSIVE)
      {
        o[k][e] = acc;
        acc = op(acc, v[k][e]);
      }
      else
      {
        acc = op(acc, v[k][e]);
        o[k][e] = acc;
      }
    }
  }
  if (full)
  {
#pragma unroll
    for (int k = 0; k < V; ++k)
    {
      lvec_t x;
#pragma unroll
      for (int e = 0; e < VEC; ++e) x[e] = o[k][e];
      __builtin_nontemporal_store(x, reinterpret_cast<lvec_t*>(out + wbase + ((int64_t)k * wave_size + lane) * VEC));
    }
  }
  else
  {
#pragma unroll
    for (int k = 0; k < V; ++k)
    {
      store_vec_guarded<S>(out, wbase + ((int64_t)k * wave_size + lane) * VEC, n, o[k]);
    }
  }
}

// Segments of at most one wave chunk (S::WCHUNK elements), one wave each:
// no descriptors, no barriers.  The element layout, per-lane folds, wave
// scans and prefix combines are those of wave 0 of scan_kernel over a
// one-tile segment, so every result is bit-identical to the tile path.
template<typename T, typename Op, bool EXCLUSIVE, int MODE>
__global__ __launch_bounds__(kBlock) void short_scan_kernel(const T* __restrict__ in,
                                                            T* __restrict__ out,
                                                            const rs::TileRec* __restrict__ recs,
                                                            int64_t nrec, T init)
{
  using S = ScanShape<T, MODE>;
  constexpr int VEC = S::VEC;
  constexpr int V = S::V;
  const Op op{};
  const int64_t ri = (int64_t)blockIdx.x * kWaves + wave_id();
  if (ri >= nrec) return;  // whole wave; the kernel has no barrier
  const rs::TileRec r = recs[ri];
  const int lane = lane_id();
  const int64_t b = r.seg_start;
  const int64_t n = r.seg_start + r.seg_len;
  T* const o_base = out + r.out_delta;

  T v[V][VEC];
#pragma unroll
  for (int k = 0; k < V; ++k)
  {
    load_vec_guarded<S>(in, b + ((int64_t)k * wave_size + lane) * VEC, n, v[k], Op::ident());
  }
  T lane_pref[V];
  T run = Op::ident();
#pragma unroll
  for (int k = 0; k < V; ++k)
  {
    T t = v[k][0];
#pragma unroll
    for (int e = 1; e < VEC; ++e) t = op(t, v[k][e]);
    T excl, total;
    (void)wave_scan(t, op, Op::ident(), excl, total);
    lane_pref[k] = op(run, excl);
    run = op(run, total);
  }
  const T start = EXCLUSIVE ? init : Op::ident();
  const T base_pref = op(start, Op::ident());
#pragma unroll
  for (int k = 0; k < V; ++k)
  {
    T acc = op(base_pref, lane_pref[k]);
    T o[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e)
    {
      if (EXCLUSIVE)
      {
        o[e] = acc;
        acc = op(acc, v[k][e]);
      }
      else
      {
        acc = op(acc, v[k][e]);
        o[e] = acc;
      }
    }
    store_vec_guarded<S>(o_base, b + ((int64_t)k * wave_size + lane) * VEC, n, o);
  }
}

// Zero lookback descriptors (used only when the stream is being captured:
// a replayed graph reuses its epoch, so each replay starts from clean ones).
__global__ void clear_desc_kernel(uint64_t* p, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
  {
    p[i] = 0ull;
  }
}

// The table shape of a segment plan: short tiles (MODE 2) when its
// segments are short on average; 8-byte values have one table shape.
template<typename T>
int table_mode(const raja_cdna4_segments_s* plan)
{
  return (sizeof(T) == 4 && plan->nseg > 0 && plan->total < plan->nseg * kShortMeanLen) ? 2 : 1;
}

// One scan launch.  tiles == nullptr: one span of n elements (ntiles
// tiles); otherwise tile table records of shape `mode` (1 or 2).  desc holds
// ntiles descriptors; with `clear` they are zeroed first (required whenever
// the memory may hold this epoch tag from an earlier use).
// Span tile of 4- and 8-byte values for a tuned vectors-per-lane choice
// (0: the default shape).  The tuned shapes are taken by aligned spans.
template<typename T>
constexpr int64_t span_tile(int v)
{
  return (int64_t)kWaves * wave_size * (v != 0 ? v : ScanShape<T>::V) * ScanShape<T>::VEC;
}

// The vectors-per-lane choices the tuner measures for a span of T (the
// default first); any other value is taken as the default.
template<typename T>
constexpr int kSpanChoices[3] = {0, sizeof(T) == 8 ? 6 : 8, sizeof(T) == 8 ? 3 : 4};

template<typename T, typename Op, bool EXCLUSIVE>
int launch_scan_core(const T* in, T* out, int64_t n, T init, hipStream_t s, const char* name,
                     const rs::TileRec* tiles, int64_t ntiles, void* desc, uint32_t epoch_tag,
                     bool clear, int mode = 1, int v = 0, bool gate_may_create = true)
{
  // one lookback launch at a time on the device (see rc::LookbackGate)
  rc::LookbackGate gate(s, gate_may_create);
  RC_TRY(gate.status());
  const bool aligned = (((uintptr_t)in | (uintptr_t)out) % 16) == 0;
  if (clear)
  {
    const int64_t words = ntiles * (int64_t)sizeof(Desc<T>) / 8;
    hipLaunchKernelGGL(clear_desc_kernel, dim3((unsigned)rc::ceil_div(words, 256 * 16) + 1),
                       dim3(256), 0, s, (uint64_t*)desc, words);
  }
  rc::TimedScope ts(name, s);
  if (tiles && mode == 2)
  {
    if constexpr (sizeof(T) == 4)
    {
      hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, false, 2>), dim3((unsigned)ntiles),
                         dim3(kBlock), 0, s, in, out, n, init, (Desc<T>*)desc, epoch_tag, tiles);
    }
    else
    {
      return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "short-segment tiles are for 4-byte values");
    }
  }
  else if (tiles)
  {
    hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, false, 1>), dim3((unsigned)ntiles),
                       dim3(kBlock), 0, s, in, out, n, init, (Desc<T>*)desc, epoch_tag, tiles);
  }
  else if (aligned && v == kSpanChoices<T>[1])
  {
    hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, true, 0, kSpanChoices<T>[1]>),
                       dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, out, n, init,
                       (Desc<T>*)desc, epoch_tag, (const rs::TileRec*)nullptr);
  }
  else if (aligned && v == kSpanChoices<T>[2])
  {
    hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, true, 0, kSpanChoices<T>[2]>),
                       dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, out, n, init,
                       (Desc<T>*)desc, epoch_tag, (const rs::TileRec*)nullptr);
  }
  else if (aligned)
  {
    hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, true, 0>), dim3((unsigned)ntiles),
                       dim3(kBlock), 0, s, in, out, n, init, (Desc<T>*)desc, epoch_tag,
                       (const rs::TileRec*)nullptr);
  }
  else
  {
    hipLaunchKernelGGL((scan_kernel<T, Op, EXCLUSIVE, false, 0>), dim3((unsigned)ntiles),
                       dim3(kBlock), 0, s, in, out, n, init, (Desc<T>*)desc, epoch_tag,
                       (const rs::TileRec*)nullptr);
  }
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}

// table == nullptr: one span of n elements; otherwise the tiles of a
// segment plan (n = total elements of the plan).  Descriptors come from the
// stream's scratch.
// A span takes the tuned shape of its (kind, dtype, length bucket) when both
// pointers are 16-byte aligned (v < 0: look it up; v >= 0: this shape).
template<typename T, typename Op, bool EXCLUSIVE>
int launch_scan(const T* in, T* out, int64_t n, T init, hipStream_t s, const char* name,
                const rs::TileTable* table = nullptr, int mode = 1, int v = -1)
{
  if (n < 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "negative length");
  if (n == 0) return RAJA_CDNA4_SUCCESS;  // empty range: no-op (pattern/scan.hpp:72-75)
  if (!in || !out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL span");
  int vv = 0;
  if (!table && (((uintptr_t)in | (uintptr_t)out) % 16) == 0)
  {
    vv = v;
    if (vv < 0)
    {
      RC_TRY(rtune::choice(EXCLUSIVE ? rtune::kExclusiveScan : rtune::kInclusiveScan,
                           rk::dtype_code<T>(), n, s, &vv));
    }
    if (vv != kSpanChoices<T>[1] && vv != kSpanChoices<T>[2]) vv = 0;
  }
  const int64_t ntiles = table ? table->ntiles : rc::ceil_div(n, span_tile<T>(vv));
  if (ntiles == 0) return RAJA_CDNA4_SUCCESS;
  if (ntiles > 0x7FFFFFFF) return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "scan too long");
  rc::StreamCtx* c = nullptr;
  RC_TRY(rc::get_ctx(s, &c));
  RC_TRY(rc::ensure_scan_desc(c, (size_t)ntiles * sizeof(Desc<T>)));
  // Descriptors carry the call's epoch; stale ones from earlier calls on this
  // stream never match.  Under graph capture the epoch is frozen into the
  // graph, so the descriptors are cleared by a kernel node instead.
  c->scan_epoch = (c->scan_epoch + 1u) & 0x3FFFFFFFu;
  if (c->scan_epoch == 0u) c->scan_epoch = 1u;
  const uint32_t epoch_tag = c->scan_epoch << 2;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RC_HIP(hipStreamIsCapturing(s, &cap));
  return launch_scan_core<T, Op, EXCLUSIVE>(
      in, out, n, init, s, name, table ? (const rs::TileRec*)table->d_tiles : nullptr, ntiles,
      c->scan_desc, epoch_tag, cap != hipStreamCaptureStatusNone, mode, vv);
}

template<typename T, bool EXCLUSIVE>
int scan_typed(int op, const void* in, void* out, int64_t n, const void* init, hipStream_t s)
{
  const char* nm = EXCLUSIVE ? "exclusive_scan" : "inclusive_scan";
  switch (op)
  {
    case RAJA_CDNA4_OP_PLUS:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpPlus<T>::ident();
      return launch_scan<T, rk::OpPlus<T>, EXCLUSIVE>((const T*)in, (T*)out, n, iv, s, nm);
    }
    case RAJA_CDNA4_OP_MIN:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMin<T>::ident();
      return launch_scan<T, rk::OpMin<T>, EXCLUSIVE>((const T*)in, (T*)out, n, iv, s, nm);
    }
    case RAJA_CDNA4_OP_MAX:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMax<T>::ident();
      return launch_scan<T, rk::OpMax<T>, EXCLUSIVE>((const T*)in, (T*)out, n, iv, s, nm);
    }
    default:
      return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "scan operator %d not supported", op);
  }
}

template<bool EXCLUSIVE>
int scan_dispatch(int dtype, int op, const void* in, void* out, int64_t n, const void* init,
                  hipStream_t s)
{
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return scan_typed<int32_t, EXCLUSIVE>(op, in, out, n, init, s);
    case RAJA_CDNA4_I64: return scan_typed<int64_t, EXCLUSIVE>(op, in, out, n, init, s);
    case RAJA_CDNA4_U32: return scan_typed<uint32_t, EXCLUSIVE>(op, in, out, n, init, s);
    case RAJA_CDNA4_U64: return scan_typed<uint64_t, EXCLUSIVE>(op, in, out, n, init, s);
    case RAJA_CDNA4_F32: return scan_typed<float, EXCLUSIVE>(op, in, out, n, init, s);
    case RAJA_CDNA4_F64: return scan_typed<double, EXCLUSIVE>(op, in, out, n, init, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}


// Segments longer than one wave chunk go through the tile kernel, the others
// through short_scan_kernel (one wave each), both on stream s.
template<typename T, typename Op, bool EXCLUSIVE, int MODE>
int seg_scan_shape(const T* in, T* out, raja_cdna4_segments_s* plan, T init, hipStream_t s,
                   const char* name)
{
  using S = ScanShape<T, MODE>;
  const rs::TileTable* tab = nullptr;
  RC_TRY(rs::get_table(plan, S::TILE, &tab, S::WCHUNK));
  if (tab->nshort > 0)
  {
    rc::TimedScope ts(name, s);
    hipLaunchKernelGGL((short_scan_kernel<T, Op, EXCLUSIVE, MODE>),
                       dim3((unsigned)rc::ceil_div(tab->nshort, kWaves)), dim3(kBlock), 0, s, in,
                       out, (const rs::TileRec*)tab->d_short, tab->nshort, init);
    RC_HIP(hipGetLastError());
  }
  return launch_scan<T, Op, EXCLUSIVE>(in, out, plan->total, init, s, name, tab, MODE);
}

template<typename T, typename Op, bool EXCLUSIVE>
int seg_scan_op(const T* in, T* out, raja_cdna4_segments_s* plan, T init, hipStream_t s,
                const char* name)
{
  if (plan->total > 0 && (!in || !out)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL span");
  if constexpr (sizeof(T) == 4)
  {
    if (table_mode<T>(plan) == 2) return seg_scan_shape<T, Op, EXCLUSIVE, 2>(in, out, plan, init, s, name);
  }
  return seg_scan_shape<T, Op, EXCLUSIVE, 1>(in, out, plan, init, s, name);
}

template<typename T, bool EXCLUSIVE>
int seg_scan_typed(int op, const void* in, void* out, raja_cdna4_segments_s* plan,
                   const void* init, hipStream_t s)
{
  const char* nm = EXCLUSIVE ? "segmented_exclusive_scan" : "segmented_inclusive_scan";
  switch (op)
  {
    case RAJA_CDNA4_OP_PLUS:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpPlus<T>::ident();
      return seg_scan_op<T, rk::OpPlus<T>, EXCLUSIVE>((const T*)in, (T*)out, plan, iv, s, nm);
    }
    case RAJA_CDNA4_OP_MIN:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMin<T>::ident();
      return seg_scan_op<T, rk::OpMin<T>, EXCLUSIVE>((const T*)in, (T*)out, plan, iv, s, nm);
    }
    case RAJA_CDNA4_OP_MAX:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMax<T>::ident();
      return seg_scan_op<T, rk::OpMax<T>, EXCLUSIVE>((const T*)in, (T*)out, plan, iv, s, nm);
    }
    default:
      return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "scan operator %d not supported", op);
  }
}

template<bool EXCLUSIVE>
int seg_scan_dispatch(int dtype, int op, const void* in, void* out, raja_cdna4_segments_s* plan,
                      const void* init, hipStream_t s)
{
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return seg_scan_typed<int32_t, EXCLUSIVE>(op, in, out, plan, init, s);
    case RAJA_CDNA4_I64: return seg_scan_typed<int64_t, EXCLUSIVE>(op, in, out, plan, init, s);
    case RAJA_CDNA4_U32: return seg_scan_typed<uint32_t, EXCLUSIVE>(op, in, out, plan, init, s);
    case RAJA_CDNA4_U64: return seg_scan_typed<uint64_t, EXCLUSIVE>(op, in, out, plan, init, s);
    case RAJA_CDNA4_F32: return seg_scan_typed<float, EXCLUSIVE>(op, in, out, plan, init, s);
    case RAJA_CDNA4_F64: return seg_scan_typed<double, EXCLUSIVE>(op, in, out, plan, init, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}

}  // namespace

namespace rscan
{
void span_choices(int dtype, int out[3])
{
  const bool wide = dtype == RAJA_CDNA4_I64 || dtype == RAJA_CDNA4_U64 || dtype == RAJA_CDNA4_F64;
  for (int i = 0; i < 3; ++i) out[i] = wide ? kSpanChoices<int64_t>[i] : kSpanChoices<int32_t>[i];
}

int tune_scan(int kind_exclusive, int dtype, const void* in, void* out, int64_t n, int v,
              hipStream_t s)
{
  switch (dtype)
  {
#define RSCAN_TUNE_CASE(CODE, T)                                                               \
  case CODE:                                                                                   \
    return kind_exclusive                                                                      \
               ? launch_scan<T, rk::OpPlus<T>, true>((const T*)in, (T*)out, n, T(0), s,         \
                                                     "exclusive_scan", nullptr, 1, v)          \
               : launch_scan<T, rk::OpPlus<T>, false>((const T*)in, (T*)out, n, T(0), s,        \
                                                      "inclusive_scan", nullptr, 1, v);
    RSCAN_TUNE_CASE(RAJA_CDNA4_I32, int32_t)
    RSCAN_TUNE_CASE(RAJA_CDNA4_I64, int64_t)
    RSCAN_TUNE_CASE(RAJA_CDNA4_U32, uint32_t)
    RSCAN_TUNE_CASE(RAJA_CDNA4_U64, uint64_t)
    RSCAN_TUNE_CASE(RAJA_CDNA4_F32, float)
    RSCAN_TUNE_CASE(RAJA_CDNA4_F64, double)
#undef RSCAN_TUNE_CASE
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
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

int launch_segmented_scan(int dtype, int op, bool exclusive, const void* in, void* out,
                          raja_cdna4_segments_s* plan, const void* init, hipStream_t s)
{
  return exclusive ? seg_scan_dispatch<true>(dtype, op, in, out, plan, init, s)
                   : seg_scan_dispatch<false>(dtype, op, in, out, plan, init, s);
}

}  // namespace rscan

namespace rws
{
size_t exclusive_i32_ws_bytes(int64_t ntiles_or_n, bool is_table)
{
  const int64_t ntiles = is_table ? ntiles_or_n : rc::ceil_div(ntiles_or_n, ScanShape<int32_t>::TILE);
  return rc::align_up((size_t)(ntiles + 1) * sizeof(Desc<int32_t>), 256);
}

int exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* ws, hipStream_t s)
{
  if (n <= 0) return RAJA_CDNA4_SUCCESS;
  const int64_t ntiles = rc::ceil_div(n, ScanShape<int32_t>::TILE);
  return launch_scan_core<int32_t, rk::OpPlus<int32_t>, true>(in, out, n, 0, s, "exclusive_scan",
                                                               nullptr, ntiles, ws, 1u << 2, true);
}

int table_exclusive_i32(const int32_t* in_base, int32_t* out_base, const rs::TileRec* d_tiles,
                        int64_t ntiles, void* ws, hipStream_t s)
{
  if (ntiles <= 0) return RAJA_CDNA4_SUCCESS;
  return launch_scan_core<int32_t, rk::OpPlus<int32_t>, true>(
      in_base, out_base, 0, 0, s, "segmented_exclusive_scan", d_tiles, ntiles, ws, 1u << 2, true);
}

int64_t scan_tile_i32() { return ScanShape<int32_t, 1>::TILE; }
}  // namespace rws

// ---- allocation-free exclusive scan ---------------------------------------------
// raja_cdna4_exclusive_scan_noalloc keeps its lookback descriptors in this
// file-scope array instead of the stream's scratch: every GPU the library
// runs on holds one copy of it until the process ends.  1 MiB holds 131072
// descriptors of 4-byte values (tiles of 32768: 2^32 elements, more than
// any int-indexed range) or 65536 of 8-byte values (tiles of 12288: about
// 8.05e8 elements).
namespace
{
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
__device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / sizeof(uint64_t)];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};  // device address of `scratch`, per GPU

// Device address of this GPU's copy of `scratch`; looked up once per device
// with that device current (the array's name on the host is only a handle).
int noalloc_scratch(int dev, void** out)
{
  if (dev < 0 || dev >= kNoAllocMaxDevices)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "device %d beyond the library's %d", dev,
                         kNoAllocMaxDevices);
  }
  std::lock_guard<std::mutex> g(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev])
  {
    void* p = nullptr;
    RC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return rc::set_error(RAJA_CDNA4_ERROR_HIP, "no device address for the scan scratch");
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return RAJA_CDNA4_SUCCESS;
}

template<typename T, typename Op>
int noalloc_scan_op(const T* in, T* out, int64_t n, T init, hipStream_t s, void* desc)
{
  const int64_t ntiles = rc::ceil_div(n, ScanShape<T>::TILE);
  if ((size_t)ntiles * sizeof(Desc<T>) > kNoAllocScratchBytes)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED,
                         "allocation-free scan of %lld elements needs %lld lookback descriptors; "
                         "its fixed scratch holds %lld",
                         (long long)n, (long long)ntiles,
                         (long long)(kNoAllocScratchBytes / sizeof(Desc<T>)));
  }
  // The descriptors keep the previous call's contents: a clear kernel (no
  // memset, so the call can also be captured into a graph) zeroes them, and
  // the fixed tag 1 then marks this call's.  No timing events are created.
  return launch_scan_core<T, Op, true>(in, out, n, init, s, nullptr, nullptr, ntiles, desc,
                                       1u << 2, true, 1, 0, /*gate_may_create=*/false);
}

template<typename T>
int noalloc_scan_typed(int op, const void* in, void* out, int64_t n, const void* init,
                       hipStream_t s, void* desc)
{
  switch (op)
  {
    case RAJA_CDNA4_OP_PLUS:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpPlus<T>::ident();
      return noalloc_scan_op<T, rk::OpPlus<T>>((const T*)in, (T*)out, n, iv, s, desc);
    }
    case RAJA_CDNA4_OP_MIN:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMin<T>::ident();
      return noalloc_scan_op<T, rk::OpMin<T>>((const T*)in, (T*)out, n, iv, s, desc);
    }
    case RAJA_CDNA4_OP_MAX:
    {
      const T iv = init ? *static_cast<const T*>(init) : rk::OpMax<T>::ident();
      return noalloc_scan_op<T, rk::OpMax<T>>((const T*)in, (T*)out, n, iv, s, desc);
    }
    default:
      return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "scan operator %d not supported", op);
  }
}
}  // namespace

namespace rscan
{
int noalloc_setup(int device) { void* p = nullptr; return noalloc_scratch(device, &p); }
}  // namespace rscan

extern "C" {

int raja_cdna4_exclusive_scan_noalloc(int dtype, int op, const void* in, void* out, int64_t n,
                                      const void* init, void* stream)
{
  if (n < 0) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "negative length");
  if (n == 0) return RAJA_CDNA4_SUCCESS;
  if (!in || !out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL span");
  int dev = 0;
  RC_HIP(hipGetDevice(&dev));
  void* desc = nullptr;
  RC_TRY(noalloc_scratch(dev, &desc));
  hipStream_t s = (hipStream_t)stream;
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return noalloc_scan_typed<int32_t>(op, in, out, n, init, s, desc);
    case RAJA_CDNA4_I64: return noalloc_scan_typed<int64_t>(op, in, out, n, init, s, desc);
    case RAJA_CDNA4_U32: return noalloc_scan_typed<uint32_t>(op, in, out, n, init, s, desc);
    case RAJA_CDNA4_U64: return noalloc_scan_typed<uint64_t>(op, in, out, n, init, s, desc);
    case RAJA_CDNA4_F32: return noalloc_scan_typed<float>(op, in, out, n, init, s, desc);
    case RAJA_CDNA4_F64: return noalloc_scan_typed<double>(op, in, out, n, init, s, desc);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}

}  // extern "C"

extern "C" {

int raja_cdna4_inclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                              void* stream)
{
  return scan_dispatch<false>(dtype, op, in, out, n, nullptr, (hipStream_t)stream);
}

int raja_cdna4_exclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                              const void* init, void* stream)
{
  return scan_dispatch<true>(dtype, op, in, out, n, init, (hipStream_t)stream);
}

}  // extern "C"
