// Shape tuning of the span scans and the radix sort, with an on-disk cache.
//
// Reference role: RAJA's HIP backend chooses launch shapes at run time from
// the occupancy calculator and its concretizers
// (include/RAJA/policy/hip/MemUtils_HIP.hpp:389-519, 547-633) and from
// compile-time tuning policies (include/RAJA/policy/hip/policy.hpp:
// 1228-1353).  Here the shapes that matter for bandwidth -- the tile of a
// span scan (vectors per lane) and the tile of a radix pass (keys per
// thread) -- are measured per (operation, element type, length bucket) on
// the device at hand and remembered in a text file keyed by device and
// runtime version.  Every candidate gives the same integer results (scans
// and sorts are exact whatever their tile), so a cold cache and a warm one
// agree bit for bit; a missing or deleted file only means default shapes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtune
{
enum Op
{
  kInclusiveScan = 0,
  kExclusiveScan = 1,
  kSortPairs = 2,
  kNumOps = 3
};

// Length bucket of n: ceil(log2 n), at least 10.
int bucket(int64_t n);

// The shape a call of (op, dtype, n) takes: 0 for the default, else the
// tuned candidate.  With RAJA_CDNA4_AUTOTUNE=1 in the environment an
// untuned bucket is measured first (never while `s` is being captured).
int choice(int op, int dtype, int64_t n, hipStream_t s, int* out);
}  // namespace rtune

// Candidate runners (scan.hip, sort.hip).
namespace rscan
{
void span_choices(int dtype, int out[3]);
int tune_scan(int kind_exclusive, int dtype, const void* in, void* out, int64_t n, int v,
              hipStream_t s);
}  // namespace rscan
namespace rsort
{
void tile_choices(int key_dtype, int out[3]);
int tune_sort(int key_dtype, void* keys, void* vals, int64_t n, int items, hipStream_t s);
}  // namespace rsort
