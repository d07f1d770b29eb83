// Launchers that take their working memory from the caller instead of the
// stream's scratch: the request server carves it from its arena, so every
// byte a request needs besides its inputs and outputs is accounted for at
// admission.  All are asynchronous on `s` and never synchronise it.
#pragma once

#include "segments.hpp"

namespace rws
{
// ---- single requests -------------------------------------------------------------
// FP64 sum of x[0..n) into *out (same tree as raja_cdna4_reduce, init 0).
size_t sum_f64_ws_bytes(int64_t n);
int sum_f64(const double* x, int64_t n, double* out, void* ws, hipStream_t s);

// int32 exclusive plus-scan, init 0.  Table form: ntiles tile records.
size_t exclusive_i32_ws_bytes(int64_t ntiles_or_n, bool is_table);
int exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* ws, hipStream_t s);

// Stable ascending sort of (uint32 key, int32 value) pairs from
// (in_keys, in_vals) into (out_keys, out_vals); every pass runs.
size_t sort_u32_i32_ws_bytes(int64_t n);
int sort_pairs_u32_i32(const uint32_t* in_keys, const int32_t* in_vals, uint32_t* out_keys,
                       int32_t* out_vals, int64_t n, void* ws, hipStream_t s);

// ---- batches of requests ---------------------------------------------------------
// Element offsets in the records below are relative to one base address
// shared by every input and output of the batch.

// Sums of nseg FP64 ranges given by tile records (tile = sum_tile_f64());
// range g's result goes to base[d_out_index[g]].  ws holds ntiles partials.
int64_t sum_tile_f64();
// Longest range table_sum_f64 folds: its final fold holds 64 groups of 64
// tile values.
int64_t sum_batch_max();
int table_sum_f64(const double* base, const rs::TileRec* d_tiles, int64_t ntiles,
                  const uint32_t* d_first, const uint32_t* d_count, int64_t nseg,
                  const int64_t* d_out_index, void* ws, hipStream_t s);

// Exclusive plus-scans (init 0) of int32 ranges given by tile records
// (tile = scan_tile_i32(); out_delta moves each range's output).
int64_t scan_tile_i32();
int table_exclusive_i32(const int32_t* in_base, int32_t* out_base, const rs::TileRec* d_tiles,
                        int64_t ntiles, void* ws, hipStream_t s);

// Stable sorts of nseg (uint32, int32) ranges of at most block_sort_max()
// pairs, one block each.
int64_t block_sort_max();
int block_sort_u32_i32(const uint32_t* base, const rs::SortSeg* d_segs, int64_t nseg, hipStream_t s);
}  // namespace rws
