// Batched (segmented) reduce, scan and sort over a segment plan.
//
// The reference runs one algorithm call per range (one forall / scan /
// sort_pairs per segment, include/RAJA/policy/hip/{reduce,scan,sort}.hpp);
// thousands of small uneven ranges then pay one launch (and for scan/sort,
// rocPRIM's temp-storage query and allocation) each.  Here one plan holds
// the device segment/tile table, and each batched call is one or two
// kernels over all segments:
//   * reduce: one block per tile of every segment, then one wave per
//     segment folds its tile values -- the same tree as the span reduction
//     (reduce_tree.hpp), so a segment's result is bit-identical alone and in
//     the batch;
//   * scan: the single-pass scan kernel over the tile table; each segment
//     starts from init and its lookback stops at its own first tile.
//     Segments of at most one wave chunk are scanned one per wave by a
//     second kernel on the same stream (no descriptors, same bits);
//   * sort_pairs: one block per segment of up to 8192 keys sorts it in LDS
//     (LSD passes with the onesweep kernel's stable wave ranking); the
//     longer segments of 4-byte keys (and 4-byte or no values) are sorted
//     together by the segmented onesweep (one histogram launch, one launch
//     per digit pass over all their tiles); 8-byte keys or values take one
//     onesweep sort per longer segment.
#include <algorithm>
#include <vector>

#include "radix.hpp"
#include "reduce_tree.hpp"
#include "segments.hpp"
#include "workspace.hpp"

namespace rs
{

int get_table(raja_cdna4_segments_s* p, int64_t tile, const TileTable** out, int64_t short_max)
{
  std::lock_guard<std::mutex> g(p->lock);
  const auto key = std::make_pair(tile, short_max);
  auto it = p->tables.find(key);
  if (it != p->tables.end())
  {
    *out = &it->second;
    return RAJA_CDNA4_SUCCESS;
  }
  TileTable t;
  t.tile = tile;
  std::vector<TileRec> recs, shorts;
  std::vector<uint32_t> first((size_t)p->nseg), count((size_t)p->nseg);
  for (int64_t s = 0; s < p->nseg; ++s)
  {
    const int64_t b = p->offsets[(size_t)s], len = p->offsets[(size_t)s + 1] - b;
    if (len > 0 && len <= short_max)
    {
      shorts.push_back(TileRec{b, len, 0u, (uint32_t)s, 0});
      first[(size_t)s] = (uint32_t)recs.size();
      count[(size_t)s] = 0;
      continue;
    }
    const int64_t nt = rc::ceil_div(len, tile);
    first[(size_t)s] = (uint32_t)recs.size();
    count[(size_t)s] = (uint32_t)nt;
    t.max_count = std::max(t.max_count, (uint32_t)nt);
    for (int64_t j = 0; j < nt; ++j) recs.push_back(TileRec{b, len, (uint32_t)j, (uint32_t)s, 0});
  }
  t.ntiles = (int64_t)recs.size();
  t.nshort = (int64_t)shorts.size();
  RC_HIP(hipMalloc(&t.d_tiles, sizeof(TileRec) * (recs.size() + 1)));
  RC_HIP(hipMalloc(&t.d_first, sizeof(uint32_t) * (first.size() + 1)));
  RC_HIP(hipMalloc(&t.d_count, sizeof(uint32_t) * (count.size() + 1)));
  RC_HIP(hipMalloc(&t.d_short, sizeof(TileRec) * (shorts.size() + 1)));
  if (!recs.empty())
  {
    RC_HIP(hipMemcpy(t.d_tiles, recs.data(), sizeof(TileRec) * recs.size(), hipMemcpyHostToDevice));
  }
  if (!shorts.empty())
  {
    RC_HIP(hipMemcpy(t.d_short, shorts.data(), sizeof(TileRec) * shorts.size(), hipMemcpyHostToDevice));
  }
  if (!first.empty())
  {
    RC_HIP(hipMemcpy(t.d_first, first.data(), sizeof(uint32_t) * first.size(), hipMemcpyHostToDevice));
    RC_HIP(hipMemcpy(t.d_count, count.data(), sizeof(uint32_t) * count.size(), hipMemcpyHostToDevice));
  }
  auto ins = p->tables.emplace(key, t);
  *out = &ins.first->second;
  return RAJA_CDNA4_SUCCESS;
}

}  // namespace rs

namespace
{
using namespace RAJA::cdna4::dev;

// ---- segmented reduce ----------------------------------------------------------

constexpr int kSegMaxGroups = 64;  // segments of up to 4096 tiles

template<typename T, typename Op>
__global__ __launch_bounds__(rt::kBlock) void seg_tiles_kernel(
    const T* __restrict__ x, const rs::TileRec* __restrict__ tiles,
    typename Op::acc_t* __restrict__ partials)
{
  using acc_t = typename Op::acc_t;
  __shared__ acc_t s_scratch[max_waves];
  const rs::TileRec r = tiles[blockIdx.x];
  const acc_t v = rt::fold_tile<T, Op>(x + r.seg_start, r.seg_len, r.j, s_scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

template<typename T, typename Op>
int seg_reduce(raja_cdna4_segments_s* p, const T* x, typename Op::acc_t init,
               typename Op::acc_t* out, hipStream_t s)
{
  using acc_t = typename Op::acc_t;
  const rs::TileTable* tab = nullptr;
  RC_TRY(rs::get_table(p, rt::tile_elems<T>(), &tab));
  if (tab->max_count > (uint32_t)kSegMaxGroups * 64)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "segment too long for the batched reduce");
  }
  if (p->nseg == 0) return RAJA_CDNA4_SUCCESS;
  rc::StreamCtx* c = nullptr;
  RC_TRY(rc::get_ctx(s, &c));
  RC_TRY(rc::ensure_tmp(c, (size_t)(tab->ntiles + 1) * sizeof(acc_t)));
  acc_t* partials = (acc_t*)c->tmp;
  rc::TimedScope ts("segmented_reduce", s);
  if (tab->ntiles > 0)
  {
    hipLaunchKernelGGL((seg_tiles_kernel<T, Op>), dim3((unsigned)tab->ntiles), dim3(rt::kBlock), 0, s,
                       x, (const rs::TileRec*)tab->d_tiles, partials);
    RC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((rt::final_kernel<Op, 1, kSegMaxGroups>), dim3((unsigned)p->nseg), dim3(64), 0, s,
                     partials, (const uint32_t*)tab->d_first, (const uint32_t*)tab->d_count,
                     (int64_t)0, init, out, (const int64_t*)nullptr);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}

template<typename T>
int seg_reduce_typed(raja_cdna4_segments_s* p, int op, const void* x, const void* init, void* out,
                     hipStream_t s)
{
  const T iv = *static_cast<const T*>(init);
  switch (op)
  {
    case RAJA_CDNA4_OP_PLUS: return seg_reduce<T, rk::OpPlus<T>>(p, (const T*)x, iv, (T*)out, s);
    case RAJA_CDNA4_OP_MIN: return seg_reduce<T, rk::OpMin<T>>(p, (const T*)x, iv, (T*)out, s);
    case RAJA_CDNA4_OP_MAX: return seg_reduce<T, rk::OpMax<T>>(p, (const T*)x, iv, (T*)out, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "operator %d not supported here", op);
  }
}

// ---- segmented sort: one block per short segment, LSD passes in LDS ----------------

template<typename K, typename V, bool HAS_VALS>
struct BlockSortShape
{
  static constexpr bool wide = sizeof(K) > 4 || (HAS_VALS && sizeof(V) > 4);
  static constexpr int BLOCK = 512;
  static constexpr int ITEMS = wide ? 8 : 16;
  static constexpr int TILE = BLOCK * ITEMS;
};

template<typename K, typename V, bool DESC, bool HAS_VALS>
__global__ __launch_bounds__(512) void block_sort_kernel(const K* in_keys, const V* in_vals,
                                                         K* out_keys, V* out_vals,
                                                         const rs::SortSeg* __restrict__ segs)
{
  using S = BlockSortShape<K, V, HAS_VALS>;
  constexpr int BLOCK = S::BLOCK;
  constexpr int ITEMS = S::ITEMS;
  constexpr int NW = BLOCK / wave_size;
  constexpr int WCHUNK = wave_size * ITEMS;
  constexpr int P = (int)sizeof(K);
  __shared__ K s_keys[S::TILE];
  __shared__ V s_vals[HAS_VALS ? S::TILE : 1];
  __shared__ uint32_t s_whist[NW * kRadix];
  __shared__ uint32_t s_tstart[kRadix];
  __shared__ uint32_t s_scan[kRadix / wave_size];

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = wave_id();
  const rs::SortSeg sg = segs[blockIdx.x];
  const int n = (int)sg.n;
  K k[ITEMS];
  V v[HAS_VALS ? ITEMS : 1];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
  {
    const int i = wave * WCHUNK + j * wave_size + lane;
    k[j] = (i < n) ? in_keys[sg.in_keys + i] : K(0);
    if (HAS_VALS) v[j] = (i < n) ? in_vals[sg.in_vals + i] : V(0);
  }
  const uint64_t lt_mask = lanemask_lt();
  for (int p = 0; p < P; ++p)
  {
    const int shift = 8 * p;
    for (int i = tid; i < NW * kRadix; i += BLOCK) s_whist[i] = 0u;
    __syncthreads();
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      const int i = wave * WCHUNK + j * wave_size + lane;
      const bool valid = i < n;
      const uint32_t d = digit_of<K, DESC>(k[j], shift);
      uint64_t peers = ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b)
      {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      const uint32_t lower = (uint32_t)__popcll(peers & lt_mask);
      uint32_t cnt = 0;
      if (valid) cnt = s_whist[wave * kRadix + d];
      rank[j] = cnt + lower;
      if (valid && lower == 0) s_whist[wave * kRadix + d] = cnt + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t ex_t = 0;
    if (tid < kRadix)
    {
      uint32_t total = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w)
      {
        const uint32_t c = s_whist[w * kRadix + tid];
        s_whist[w * kRadix + tid] = total;
        total += c;
      }
      uint32_t tot;
      (void)wave_scan(total, rk::OpPlus<uint32_t>{}, 0u, ex_t, tot);
      if (lane == 0) s_scan[wave] = tot;
    }
    __syncthreads();
    if (tid < kRadix)
    {
      uint32_t ts = ex_t;
      for (int w = 0; w < wave; ++w) ts += s_scan[w];
      s_tstart[tid] = ts;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      const int i = wave * WCHUNK + j * wave_size + lane;
      if (i < n)
      {
        const uint32_t d = digit_of<K, DESC>(k[j], shift);
        const uint32_t pos = s_tstart[d] + s_whist[wave * kRadix + d] + rank[j];
        s_keys[pos] = k[j];
        if (HAS_VALS) s_vals[pos] = v[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
      const int i = wave * WCHUNK + j * wave_size + lane;
      if (i < n)
      {
        k[j] = s_keys[i];
        if (HAS_VALS) v[j] = s_vals[i];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
  {
    const int i = wave * WCHUNK + j * wave_size + lane;
    if (i < n)
    {
      out_keys[sg.out_keys + i] = k[j];
      if (HAS_VALS) out_vals[sg.out_vals + i] = v[j];
    }
  }
}

template<typename K, typename V, bool HAS_VALS, bool DESC>
int seg_sort(raja_cdna4_segments_s* p, K* keys, V* vals, int key_dtype, int val_bytes,
             hipStream_t s)
{
  using S = BlockSortShape<K, V, HAS_VALS>;
  std::vector<rs::SortSeg> small;
  std::vector<int64_t> big;
  for (int64_t g = 0; g < p->nseg; ++g)
  {
    const int64_t b = p->offsets[(size_t)g];
    const int64_t len = p->offsets[(size_t)g + 1] - b;
    if (len <= 1) continue;
    if (len <= S::TILE) small.push_back(rs::SortSeg{b, b, b, b, len});
    else big.push_back(g);
  }
  if (!small.empty())
  {
    rc::StreamCtx* c = nullptr;
    RC_TRY(rc::get_ctx(s, &c));
    // short-segment records: uploaded once per plan and block shape
    rs::SortSeg* d_segs = nullptr;
    {
      std::lock_guard<std::mutex> g(p->lock);
      auto it = p->sort_segs.find((int64_t)S::TILE);
      if (it == p->sort_segs.end())
      {
        RC_HIP(hipMalloc(&d_segs, sizeof(rs::SortSeg) * small.size()));
        RC_HIP(hipMemcpy(d_segs, small.data(), sizeof(rs::SortSeg) * small.size(), hipMemcpyHostToDevice));
        it = p->sort_segs.emplace((int64_t)S::TILE, d_segs).first;
      }
      d_segs = it->second;
    }
    rc::TimedScope ts("segmented_block_sort", s);
    hipLaunchKernelGGL((block_sort_kernel<K, V, DESC, HAS_VALS>), dim3((unsigned)small.size()),
                       dim3(S::BLOCK), 0, s, keys, vals, keys, vals, (const rs::SortSeg*)d_segs);
    RC_HIP(hipGetLastError());
  }
  if (big.empty()) return RAJA_CDNA4_SUCCESS;
  if constexpr (sizeof(K) == 4 && (!HAS_VALS || sizeof(V) == 4))
  {
    // Every longer segment in one segmented onesweep: tile records built
    // once per plan; alternates, histograms and descriptors from the
    // stream's scratch.
    const int64_t tile = rsort::seg_tile();
    raja_cdna4_segments_s::SortTable tab;
    {
      std::lock_guard<std::mutex> g(p->lock);
      auto it = p->sort_tables.find((int64_t)S::TILE);
      if (it == p->sort_tables.end())
      {
        std::vector<rs::SortTile> recs;
        for (size_t r = 0; r < big.size(); ++r)
        {
          const int64_t b = p->offsets[(size_t)big[r]];
          const int64_t len = p->offsets[(size_t)big[r] + 1] - b;
          for (int64_t j = 0; j * tile < len; ++j)
          {
            recs.push_back(rs::SortTile{b, b, b, b, b, b, len, (uint32_t)j, (uint32_t)r});
          }
        }
        raja_cdna4_segments_s::SortTable t;
        t.ntiles = (int64_t)recs.size();
        t.nranges = (int64_t)big.size();
        RC_HIP(hipMalloc(&t.d_tiles, sizeof(rs::SortTile) * recs.size()));
        RC_HIP(hipMemcpy(t.d_tiles, recs.data(), sizeof(rs::SortTile) * recs.size(),
                         hipMemcpyHostToDevice));
        it = p->sort_tables.emplace((int64_t)S::TILE, t).first;
      }
      tab = it->second;
    }
    rc::StreamCtx* c = nullptr;
    RC_TRY(rc::get_ctx(s, &c));
    const size_t kb = rc::align_up((size_t)p->total * sizeof(K), 256);
    const size_t vb = HAS_VALS ? rc::align_up((size_t)p->total * sizeof(V), 256) : 0;
    RC_TRY(rc::ensure_tmp(c, kb + vb + rsort::seg_hist_bytes(tab.nranges)));
    RC_TRY(rc::ensure_sort_desc(c, rsort::seg_desc_bytes(tab.ntiles)));
    char* tmp = (char*)c->tmp;
    return rsort::seg_sort_table(key_dtype, HAS_VALS ? 4 : 0, DESC ? 1 : 0, keys, vals, keys, vals,
                                 tmp, HAS_VALS ? (void*)(tmp + kb) : nullptr, tab.d_tiles,
                                 tab.ntiles, tab.nranges, (uint32_t*)(tmp + kb + vb),
                                 (uint32_t*)c->sort_desc, s);
  }
  else
  {
    // 8-byte keys or values: one onesweep sort per longer segment.
    for (int64_t g : big)
    {
      const int64_t b = p->offsets[(size_t)g], len = p->offsets[(size_t)g + 1] - b;
      RC_TRY(raja_cdna4_sort_pairs(key_dtype, keys + b, HAS_VALS ? (void*)(vals + b) : nullptr,
                                   val_bytes, len, DESC ? 1 : 0, s));
    }
    return RAJA_CDNA4_SUCCESS;
  }
}

template<typename K>
int seg_sort_key(raja_cdna4_segments_s* p, int kd, void* keys, void* vals, int vb, int desc,
                 hipStream_t s)
{
  if (!vals)
  {
    return desc ? seg_sort<K, uint32_t, false, true>(p, (K*)keys, nullptr, kd, 0, s)
                : seg_sort<K, uint32_t, false, false>(p, (K*)keys, nullptr, kd, 0, s);
  }
  if (vb == 4)
  {
    return desc ? seg_sort<K, uint32_t, true, true>(p, (K*)keys, (uint32_t*)vals, kd, 4, s)
                : seg_sort<K, uint32_t, true, false>(p, (K*)keys, (uint32_t*)vals, kd, 4, s);
  }
  if (vb == 8)
  {
    return desc ? seg_sort<K, uint64_t, true, true>(p, (K*)keys, (uint64_t*)vals, kd, 8, s)
                : seg_sort<K, uint64_t, true, false>(p, (K*)keys, (uint64_t*)vals, kd, 8, s);
  }
  return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "value size %d (4 or 8 supported)", vb);
}

}  // namespace

namespace rws
{
int64_t sum_tile_f64() { return rt::tile_elems<double>(); }

int64_t sum_batch_max() { return (int64_t)kSegMaxGroups * 64 * rt::tile_elems<double>(); }

int table_sum_f64(const double* base, const rs::TileRec* d_tiles, int64_t ntiles,
                  const uint32_t* d_first, const uint32_t* d_count, int64_t nseg,
                  const int64_t* d_out_index, void* ws, hipStream_t s)
{
  using Op = rk::OpPlus<double>;
  if (nseg == 0) return RAJA_CDNA4_SUCCESS;
  double* partials = (double*)ws;
  rc::TimedScope ts("segmented_reduce", s);
  if (ntiles > 0)
  {
    hipLaunchKernelGGL((seg_tiles_kernel<double, Op>), dim3((unsigned)ntiles), dim3(rt::kBlock), 0, s,
                       base, d_tiles, partials);
    RC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((rt::final_kernel<Op, 1, kSegMaxGroups>), dim3((unsigned)nseg), dim3(64), 0, s,
                     (const double*)partials, d_first, d_count, (int64_t)0, 0.0, (double*)base,
                     d_out_index);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}

int64_t block_sort_max() { return BlockSortShape<uint32_t, uint32_t, true>::TILE; }

int block_sort_u32_i32(const uint32_t* base, const rs::SortSeg* d_segs, int64_t nseg, hipStream_t s)
{
  if (nseg == 0) return RAJA_CDNA4_SUCCESS;
  rc::TimedScope ts("segmented_block_sort", s);
  hipLaunchKernelGGL((block_sort_kernel<uint32_t, uint32_t, false, true>), dim3((unsigned)nseg),
                     dim3(512), 0, s, base, base, (uint32_t*)base, (uint32_t*)base, d_segs);
  RC_HIP(hipGetLastError());
  return RAJA_CDNA4_SUCCESS;
}
}  // namespace rws

extern "C" {

int raja_cdna4_segments_create(const int64_t* offsets, int64_t nseg, raja_cdna4_segments_s** out)
{
  if (!out || nseg < 0 || (nseg > 0 && !offsets))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad segment arguments");
  }
  if (nseg > 0x7FFFFFFF) return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "too many segments");
  for (int64_t s = 0; s < nseg; ++s)
  {
    if (offsets[s + 1] < offsets[s] || offsets[0] < 0)
    {
      return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    }
  }
  raja_cdna4_segments_s* p = new raja_cdna4_segments_s();
  RC_HIP(hipGetDevice(&p->device));
  p->nseg = nseg;
  p->offsets.assign(offsets, offsets + nseg + 1);
  p->total = p->offsets.back();
  if (hipMalloc(&p->d_offsets, sizeof(int64_t) * (size_t)(nseg + 1)) != hipSuccess ||
      hipMemcpy(p->d_offsets, offsets, sizeof(int64_t) * (size_t)(nseg + 1), hipMemcpyHostToDevice) !=
          hipSuccess)
  {
    (void)hipGetLastError();
    delete p;
    return rc::set_error(RAJA_CDNA4_ERROR_OUT_OF_MEMORY, "segment plan upload failed");
  }
  *out = p;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_segments_destroy(raja_cdna4_segments_s* p)
{
  if (!p) return RAJA_CDNA4_SUCCESS;
  RC_HIP(hipDeviceSynchronize());
  for (auto& kv : p->tables)
  {
    if (kv.second.d_tiles) RC_HIP(hipFree(kv.second.d_tiles));
    if (kv.second.d_first) RC_HIP(hipFree(kv.second.d_first));
    if (kv.second.d_count) RC_HIP(hipFree(kv.second.d_count));
    if (kv.second.d_short) RC_HIP(hipFree(kv.second.d_short));
  }
  for (auto& kv : p->sort_segs) RC_HIP(hipFree(kv.second));
  for (auto& kv : p->sort_tables) RC_HIP(hipFree(kv.second.d_tiles));
  if (p->d_offsets) RC_HIP(hipFree(p->d_offsets));
  delete p;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_segmented_reduce(raja_cdna4_segments_s* p, int dtype, int op, const void* x,
                                const void* init, void* out_dev, void* stream)
{
  if (!p || !init || !out_dev || (p->total > 0 && !x))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL argument");
  }
  hipStream_t s = (hipStream_t)stream;
  switch (dtype)
  {
    case RAJA_CDNA4_I32: return seg_reduce_typed<int32_t>(p, op, x, init, out_dev, s);
    case RAJA_CDNA4_I64: return seg_reduce_typed<int64_t>(p, op, x, init, out_dev, s);
    case RAJA_CDNA4_F32: return seg_reduce_typed<float>(p, op, x, init, out_dev, s);
    case RAJA_CDNA4_F64: return seg_reduce_typed<double>(p, op, x, init, out_dev, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "dtype %d not supported", dtype);
  }
}

int raja_cdna4_segmented_exclusive_scan(raja_cdna4_segments_s* p, int dtype, int op, const void* in,
                                        void* out, const void* init, void* stream)
{
  if (!p) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL plan");
  return rscan::launch_segmented_scan(dtype, op, true, in, out, p, init, (hipStream_t)stream);
}

int raja_cdna4_segmented_inclusive_scan(raja_cdna4_segments_s* p, int dtype, int op, const void* in,
                                        void* out, void* stream)
{
  if (!p) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL plan");
  return rscan::launch_segmented_scan(dtype, op, false, in, out, p, nullptr, (hipStream_t)stream);
}

int raja_cdna4_segmented_sort_pairs(raja_cdna4_segments_s* p, int key_dtype, void* keys, void* vals,
                                    int val_bytes, int descending, void* stream)
{
  if (!p || (p->total > 0 && !keys)) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  switch (key_dtype)
  {
    case RAJA_CDNA4_U32: return seg_sort_key<uint32_t>(p, key_dtype, keys, vals, val_bytes, descending, s);
    case RAJA_CDNA4_I32: return seg_sort_key<int32_t>(p, key_dtype, keys, vals, val_bytes, descending, s);
    case RAJA_CDNA4_F32: return seg_sort_key<float>(p, key_dtype, keys, vals, val_bytes, descending, s);
    case RAJA_CDNA4_U64: return seg_sort_key<uint64_t>(p, key_dtype, keys, vals, val_bytes, descending, s);
    case RAJA_CDNA4_I64: return seg_sort_key<int64_t>(p, key_dtype, keys, vals, val_bytes, descending, s);
    case RAJA_CDNA4_F64: return seg_sort_key<double>(p, key_dtype, keys, vals, val_bytes, descending, s);
    default: return rc::set_error(RAJA_CDNA4_ERROR_UNSUPPORTED, "key dtype %d", key_dtype);
  }
}

}  // extern "C"
