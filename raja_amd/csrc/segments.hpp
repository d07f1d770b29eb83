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

// One short segment sorted by one block: where its keys and values are read
// and written (element offsets from the kernel's base pointers), and its
// length.
struct SortSeg
{
  int64_t in_keys;
  int64_t in_vals;
  int64_t out_keys;
  int64_t out_vals;
  int64_t n;
};

// One tile of one range sorted by the segmented onesweep: element offsets of
// the range's keys and values in the input (ik, iv), output (ok, ov) and
// alternate (ak, av) buffers, each from that buffer's base pointer; the
// range length, the tile's index inside the range and the range's index
// (its digit-histogram row).
struct SortTile
{
  int64_t ik, iv, ok, ov, ak, av;
  int64_t len;
  uint32_t j;
  uint32_t seg;
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
// and tile count.  A table built with a short-segment limit holds the tiles
// of the longer segments only, and the short ones (1 .. limit elements) as
// a separate list of one record each.
struct TileTable
{
  int64_t tile = 0;
  int64_t ntiles = 0;
  TileRec* d_tiles = nullptr;
  uint32_t* d_first = nullptr;
  uint32_t* d_count = nullptr;
  uint32_t max_count = 0;
  int64_t nshort = 0;
  TileRec* d_short = nullptr;
};

}  // namespace rs

struct raja_cdna4_segments_s
{
  int device = 0;
  int64_t nseg = 0;
  int64_t total = 0;
  std::vector<int64_t> offsets;  // nseg + 1, host copy
  int64_t* d_offsets = nullptr;
  std::map<std::pair<int64_t, int64_t>, rs::TileTable> tables;  // (tile, short limit)
  std::map<int64_t, rs::SortSeg*> sort_segs;  // per block shape: short-segment records
  // tile records of the segments longer than one block sort (segmented
  // onesweep), per block-sort limit
  struct SortTable
  {
    rs::SortTile* d_tiles = nullptr;
    int64_t ntiles = 0;
    int64_t nranges = 0;
  };
  std::map<int64_t, SortTable> sort_tables;
  std::mutex lock;
};

namespace rs
{
// Table for tile size `tile` (built on first use, synchronously); with
// short_max > 0, segments of at most short_max elements go to the short list.
int get_table(raja_cdna4_segments_s* p, int64_t tile, const TileTable** out, int64_t short_max = 0);
}  // namespace rs

// Segmented onesweep (sort.hip): every range of a tile table of SortTile
// records (tile = seg_tile() keys, 4-byte keys, values of 4 bytes or none)
// sorted in one histogram launch and one launch per digit pass, from the
// input through the alternates into the output buffers.  hist holds
// seg_hist_bytes(nranges), desc seg_desc_bytes(ntiles); neither needs
// clearing.
namespace rsort
{
int64_t seg_tile();
size_t seg_desc_bytes(int64_t ntiles);
size_t seg_hist_bytes(int64_t nranges);
int seg_sort_table(int key_dtype, int val_bytes, int descending, const void* in_k, const void* in_v,
                   void* out_k, void* out_v, void* alt_k, void* alt_v, const rs::SortTile* d_recs,
                   int64_t ntiles, int64_t nranges, uint32_t* hist, uint32_t* desc, hipStream_t s);
}  // namespace rsort

// Launcher implemented in scan.hip for segmented scans (SEG mode of the
// single-pass scan kernel; lookback stops at each segment's first tile).
namespace rscan
{
int launch_segmented_scan(int dtype, int op, bool exclusive, const void* in, void* out,
                          raja_cdna4_segments_s* plan, const void* init, hipStream_t s);
}  // namespace rscan
