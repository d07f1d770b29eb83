// Segment plans: the device segment/tile table behind the batched entry
// points.  A plan is built once from host offsets and reused by every
// segmented reduce/scan/sort call on it (and by graph replays of them).
#pragma once

#include <map>
#include <mutex>
#include <vector>

#include "runtime.hpp"

namespace rs
{

// One tile of one segment.
struct TileRec
{
  int64_t seg_start;  // element offset of the segment in the input
  int64_t seg_len;    // segment length
  uint32_t j;         // tile index inside the segment
  uint32_t seg;       // segment index
  int64_t out_delta;  // output element offset minus input element offset
};

// One short segment sorted by one block: input start, length, and where its
// sorted keys and values go.
struct SortSeg
{
  int64_t in_start;
  int64_t n;
  int64_t out_keys;
  int64_t out_vals;
};

// Tile records of segments given as (input start, length, output start).
inline void append_tiles(std::vector<TileRec>& recs, int64_t start, int64_t len, int64_t out_start,
                         uint32_t seg, int64_t tile)
{
  const int64_t nt = (len + tile - 1) / tile;
  for (int64_t j = 0; j < nt; ++j)
  {
    recs.push_back(TileRec{start, len, (uint32_t)j, seg, out_start - start});
  }
}

// Tiles of one tile size: tile records plus, per segment, its first tile
// and tile count.
struct TileTable
{
  int64_t tile = 0;
  int64_t ntiles = 0;
  TileRec* d_tiles = nullptr;
  uint32_t* d_first = nullptr;
  uint32_t* d_count = nullptr;
  uint32_t max_count = 0;
};

}  // namespace rs

struct raja_cdna4_segments_s
{
  int device = 0;
  int64_t nseg = 0;
  int64_t total = 0;
  std::vector<int64_t> offsets;  // nseg + 1, host copy
  int64_t* d_offsets = nullptr;
  std::map<int64_t, rs::TileTable> tables;
  std::map<int64_t, rs::SortSeg*> sort_segs;  // per block shape: short-segment records
  std::mutex lock;
};

namespace rs
{
// Table for tile size `tile` (built on first use, synchronously).
int get_table(raja_cdna4_segments_s* p, int64_t tile, const TileTable** out);
}  // namespace rs

// Launchers implemented in scan.hip for segmented scans (SEG mode of the
// single-pass scan kernel; lookback stops at each segment's first tile).
namespace rscan
{
int launch_segmented_scan(int dtype, int op, bool exclusive, const void* in, void* out,
                          raja_cdna4_segments_s* plan, const void* init, hipStream_t s);
// Plain int32 exclusive plus-scan over an uploaded tile table (serving path).
int launch_table_exclusive_scan_i32(const int32_t* in, int32_t* out, const rs::TileTable& table,
                                    int64_t total, int32_t init, hipStream_t s);
}  // namespace rscan

// Launchers implemented in segmented.hip for the serving path.
namespace rseg
{
// Sum of every segment of FP64 x (same tree as raja_cdna4_reduce).
// Segment s's result goes to out[d_out_index[s]] (out[s] when NULL).
int launch_table_sum_f64(const double* x, const rs::TileTable& table, int64_t nseg, double* out,
                         const int64_t* d_out_index, hipStream_t s);
// Stable sort of short segments (<= 8192 pairs) in LDS, one block each,
// from (in_keys, in_vals) to (out_keys, out_vals) as the records say.
int launch_block_sort_u32_i32(const uint32_t* in_keys, const int32_t* in_vals, uint32_t* out_keys,
                              int32_t* out_vals, const rs::SortSeg* d_segs, int64_t nsmall,
                              hipStream_t s);
}  // namespace rseg

namespace rsort
{
// onesweep sort_pairs without the host read-back of the histograms (every
// pass runs), so the calling thread never blocks.
int sort_pairs_u32_i32_async(uint32_t* keys, int32_t* vals, int64_t n, hipStream_t s);
}  // namespace rsort
