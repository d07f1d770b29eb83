// Request server: arena-backed admission, coalescing of small requests into
// batched launches, large requests alone on round-robin streams.
//
// The reference serves a request with one algorithm call, whose temporaries
// come from a pool (include/RAJA/util/basic_mempool.hpp:54-454) and whose
// launches queue on one stream; many small calls are then bound by launch
// latency.  Here:
//   * the arena is one device allocation carved first-fit on the host; a
//     unit (batch or single request) holds its working memory -- tables,
//     partials, lookback descriptors, sort alternates -- from launch to
//     retirement, so the memory in use is always accounted for;
//   * a batch takes the waiting small requests in arrival order and runs
//     each kind with one table upload and one or two kernels: sums over the
//     tile table of the canonical reduction tree (results land in each
//     request's own output), scans with the segmented single-pass scan
//     (tile records carry each request's output offset), sorts with one
//     block per request in LDS;
//   * completion is observed by polling one event per unit, so new arrivals
//     are admitted while earlier units run.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "workspace.hpp"

namespace
{

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;

// First-fit allocator over [0, size) with coalescing of freed ranges.
struct Arena
{
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_ranges;  // offset -> length

  void reset()
  {
    free_ranges.clear();
    if (size) free_ranges.emplace(0, size);
  }
  bool alloc(size_t bytes, size_t* off)
  {
    bytes = rc::align_up(bytes ? bytes : 1, kAlign);
    for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it)
    {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t next = it->first + bytes;
      free_ranges.erase(it);
      if (rest) free_ranges.emplace(next, rest);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes)
  {
    bytes = rc::align_up(bytes ? bytes : 1, kAlign);
    auto it = free_ranges.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_ranges.end() && it->first + it->second == nx->first)
    {
      it->second += nx->second;
      free_ranges.erase(nx);
    }
    if (it != free_ranges.begin())
    {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first)
      {
        pv->second += it->second;
        free_ranges.erase(it);
      }
    }
  }
};

// A unit slot: the pinned staging of a batch's tables and the completion
// event.  A slot is reused only after its unit retired.
struct Slot
{
  hipEvent_t done = nullptr;
  char* staging = nullptr;
  size_t staging_cap = 0;
};

struct Unit
{
  int slot = -1;
  int id = 0;
  size_t off = 0, bytes = 0;
  std::vector<int64_t> reqs;
};

bool is_batchable(const raja_cdna4_serve_config& cfg, const raja_cdna4_request& r)
{
  if (cfg.mode != RAJA_CDNA4_SERVE_BATCHED) return false;
  if (r.kind == RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32) return r.n <= rws::block_sort_max();
  return r.n <= cfg.batch_limit;
}

// Working memory of a request that runs alone.
size_t single_ws(const raja_cdna4_request& r)
{
  switch (r.kind)
  {
    case RAJA_CDNA4_REQ_SUM_F64: return rws::sum_f64_ws_bytes(r.n);
    case RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32: return rws::exclusive_i32_ws_bytes(r.n, false);
    default: return rws::sort_u32_i32_ws_bytes(r.n);
  }
}

// Upper bound of the bytes one batchable request adds to a batch's device
// image (records, partials, descriptors).
size_t batch_bytes_bound(const raja_cdna4_request& r)
{
  switch (r.kind)
  {
    case RAJA_CDNA4_REQ_SUM_F64:
    {
      const size_t t = (size_t)rc::ceil_div(r.n, rws::sum_tile_f64());
      return t * (sizeof(rs::TileRec) + sizeof(double)) + 2 * sizeof(uint32_t) + sizeof(int64_t);
    }
    case RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32:
    {
      const size_t t = (size_t)rc::ceil_div(r.n, rws::scan_tile_i32());
      return t * (sizeof(rs::TileRec) + 16);
    }
    default: return sizeof(rs::SortSeg);
  }
}
constexpr size_t kBatchFixedBound = 6 * 16 + 4 * kAlign + 16;

// Host image of one batch: the records of each kind plus the layout of the
// device copy (records first, then partials and descriptors).
struct BatchPlan
{
  std::vector<rs::TileRec> sum_tiles, scan_tiles;
  std::vector<uint32_t> sum_first, sum_count;
  std::vector<int64_t> sum_out;
  std::vector<rs::SortSeg> sorts;
  uintptr_t sum_base = 0, scan_base = 0, sort_base = 0;
  size_t o_sum_tiles = 0, o_sum_first = 0, o_sum_count = 0, o_sum_out = 0, o_scan_tiles = 0,
         o_sorts = 0, rec_bytes = 0, o_partials = 0, o_desc = 0, total = 0;

  void clear()
  {
    sum_tiles.clear();
    scan_tiles.clear();
    sum_first.clear();
    sum_count.clear();
    sum_out.clear();
    sorts.clear();
  }

  static uintptr_t lowest(uintptr_t a, const void* p)
  {
    const uintptr_t v = (uintptr_t)p;
    return (a == 0 || v < a) ? v : a;
  }

  // Builds the records of the requests `ids` (all batchable).
  void build(const raja_cdna4_request* req, const std::vector<int64_t>& ids)
  {
    clear();
    sum_base = scan_base = sort_base = 0;
    for (int64_t i : ids)
    {
      const raja_cdna4_request& r = req[i];
      if (r.kind == RAJA_CDNA4_REQ_SUM_F64)
      {
        sum_base = lowest(lowest(sum_base, r.out), r.n ? r.in : r.out);
      }
      else if (r.kind == RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32)
      {
        if (r.n) scan_base = lowest(lowest(scan_base, r.in), r.out);
      }
      else if (r.n)
      {
        sort_base = lowest(lowest(lowest(lowest(sort_base, r.in), r.in_vals), r.out), r.out_vals);
      }
    }
    const int64_t sum_tile = rws::sum_tile_f64();
    const int64_t scan_tile = rws::scan_tile_i32();
    for (int64_t i : ids)
    {
      const raja_cdna4_request& r = req[i];
      if (r.kind == RAJA_CDNA4_REQ_SUM_F64)
      {
        const int64_t start = r.n ? (int64_t)(((uintptr_t)r.in - sum_base) / 8) : 0;
        sum_first.push_back((uint32_t)sum_tiles.size());
        const uint32_t seg = (uint32_t)sum_out.size();
        rs::append_tiles(sum_tiles, start, r.n, start, seg, sum_tile);
        sum_count.push_back((uint32_t)(sum_tiles.size() - sum_first.back()));
        sum_out.push_back((int64_t)(((uintptr_t)r.out - sum_base) / 8));
      }
      else if (r.kind == RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32)
      {
        if (!r.n) continue;
        const int64_t start = (int64_t)(((uintptr_t)r.in - scan_base) / 4);
        const int64_t ostart = (int64_t)(((uintptr_t)r.out - scan_base) / 4);
        rs::append_tiles(scan_tiles, start, r.n, ostart, (uint32_t)(scan_tiles.size()), scan_tile);
      }
      else if (r.n)
      {
        sorts.push_back(rs::SortSeg{(int64_t)(((uintptr_t)r.in - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.in_vals - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.out - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.out_vals - sort_base) / 4), r.n});
      }
    }
    size_t o = 0;
    auto place = [&](size_t bytes) {
      const size_t at = o;
      o = rc::align_up(o + bytes, 16);
      return at;
    };
    o_sum_tiles = place(sizeof(rs::TileRec) * sum_tiles.size());
    o_scan_tiles = place(sizeof(rs::TileRec) * scan_tiles.size());
    o_sorts = place(sizeof(rs::SortSeg) * sorts.size());
    o_sum_out = place(sizeof(int64_t) * sum_out.size());
    o_sum_first = place(sizeof(uint32_t) * sum_first.size());
    o_sum_count = place(sizeof(uint32_t) * sum_count.size());
    rec_bytes = o;
    o = rc::align_up(o, kAlign);
    o_partials = o;
    o += rc::align_up(sizeof(double) * (sum_tiles.size() + 1), kAlign);
    o_desc = o;
    o += rws::exclusive_i32_ws_bytes((int64_t)scan_tiles.size(), true);
    total = o;
  }

  void write(char* dst) const
  {
    ::memcpy(dst + o_sum_tiles, sum_tiles.data(), sizeof(rs::TileRec) * sum_tiles.size());
    ::memcpy(dst + o_scan_tiles, scan_tiles.data(), sizeof(rs::TileRec) * scan_tiles.size());
    ::memcpy(dst + o_sorts, sorts.data(), sizeof(rs::SortSeg) * sorts.size());
    ::memcpy(dst + o_sum_out, sum_out.data(), sizeof(int64_t) * sum_out.size());
    ::memcpy(dst + o_sum_first, sum_first.data(), sizeof(uint32_t) * sum_first.size());
    ::memcpy(dst + o_sum_count, sum_count.data(), sizeof(uint32_t) * sum_count.size());
  }
};

int check_request(const raja_cdna4_request& r)
{
  if (r.kind < RAJA_CDNA4_REQ_SUM_F64 || r.kind > RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "unknown request kind %d", r.kind);
  }
  if (r.n < 0 || r.n > 0x7FFFFFFF)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "request length %lld out of range",
                         (long long)r.n);
  }
  const size_t a = r.kind == RAJA_CDNA4_REQ_SUM_F64 ? 8 : 4;
  const bool need_out = r.kind == RAJA_CDNA4_REQ_SUM_F64 || r.n > 0;
  if ((need_out && (!r.out || (uintptr_t)r.out % a)) || (r.n > 0 && (!r.in || (uintptr_t)r.in % a)))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "request buffers NULL or misaligned");
  }
  if (r.kind == RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32 && r.n > 0 &&
      (!r.in_vals || !r.out_vals || (uintptr_t)r.in_vals % 4 || (uintptr_t)r.out_vals % 4))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "sort values NULL or misaligned");
  }
  return RAJA_CDNA4_SUCCESS;
}

}  // namespace

struct raja_cdna4_server_s
{
  raja_cdna4_serve_config cfg{};
  int device = 0;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<Slot> slots;
  BatchPlan plan;
};

namespace
{

int launch_single(const raja_cdna4_request& r, char* ws, hipStream_t s)
{
  switch (r.kind)
  {
    case RAJA_CDNA4_REQ_SUM_F64:
      return rws::sum_f64((const double*)r.in, r.n, (double*)r.out, ws, s);
    case RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32:
      return rws::exclusive_i32((const int32_t*)r.in, (int32_t*)r.out, r.n, ws, s);
    default:
      return rws::sort_pairs_u32_i32((const uint32_t*)r.in, (const int32_t*)r.in_vals,
                                     (uint32_t*)r.out, (int32_t*)r.out_vals, r.n, ws, s);
  }
}

int launch_batch(raja_cdna4_server_s* sv, Slot& slot, char* dev, hipStream_t s)
{
  const BatchPlan& p = sv->plan;
  if (p.rec_bytes)
  {
    if (slot.staging_cap < p.rec_bytes)
    {
      if (slot.staging) RC_HIP(hipHostFree(slot.staging));
      slot.staging = nullptr;
      slot.staging_cap = 0;
      const size_t cap = rc::align_up(p.rec_bytes * 2, 4096);
      RC_HIP(hipHostMalloc((void**)&slot.staging, cap, hipHostMallocDefault));
      slot.staging_cap = cap;
    }
    p.write(slot.staging);
    RC_HIP(hipMemcpyAsync(dev, slot.staging, p.rec_bytes, hipMemcpyHostToDevice, s));
  }
  if (!p.sum_out.empty())
  {
    RC_TRY(rws::table_sum_f64((const double*)p.sum_base, (const rs::TileRec*)(dev + p.o_sum_tiles),
                              (int64_t)p.sum_tiles.size(), (const uint32_t*)(dev + p.o_sum_first),
                              (const uint32_t*)(dev + p.o_sum_count), (int64_t)p.sum_out.size(),
                              (const int64_t*)(dev + p.o_sum_out), dev + p.o_partials, s));
  }
  if (!p.scan_tiles.empty())
  {
    RC_TRY(rws::table_exclusive_i32((const int32_t*)p.scan_base, (int32_t*)p.scan_base,
                                    (const rs::TileRec*)(dev + p.o_scan_tiles),
                                    (int64_t)p.scan_tiles.size(), dev + p.o_desc, s));
  }
  if (!p.sorts.empty())
  {
    RC_TRY(rws::block_sort_u32_i32((const uint32_t*)p.sort_base, (const rs::SortSeg*)(dev + p.o_sorts),
                                   (int64_t)p.sorts.size(), s));
  }
  return RAJA_CDNA4_SUCCESS;
}

double seconds_since(Clock::time_point t0)
{
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Waits for every unit in flight (error path) and returns their memory.
void drain(raja_cdna4_server_s* sv, std::deque<Unit>& inflight)
{
  for (hipStream_t s : sv->streams) (void)hipStreamSynchronize(s);
  for (Unit& u : inflight) sv->arena.release(u.off, u.bytes);
  inflight.clear();
}

int run(raja_cdna4_server_s* sv, const raja_cdna4_request* req, int64_t nreq,
        raja_cdna4_result* res, double* wall_s)
{
  const raja_cdna4_serve_config& cfg = sv->cfg;
  for (int64_t i = 0; i < nreq; ++i)
  {
    RC_TRY(check_request(req[i]));
    if (i > 0 && req[i].arrival_s < req[i - 1].arrival_s)
    {
      return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "arrival times must be non-decreasing");
    }
    res[i] = raja_cdna4_result{RAJA_CDNA4_SUCCESS, -1, 0, 0, 0.0, 0.0};
  }
  RC_HIP(hipSetDevice(sv->device));
  sv->arena.reset();
  std::deque<Unit> inflight;
  std::vector<int> free_slots;
  for (int k = (int)sv->slots.size() - 1; k >= 0; --k) free_slots.push_back(k);
  const int nstreams = cfg.mode == RAJA_CDNA4_SERVE_BATCHED ? (int)sv->streams.size() : 1;
  int next_stream = 0;
  int units = 0;
  int64_t next = 0, completed = 0;
  double last_done = 0.0;
  const Clock::time_point t0 = Clock::now();
  std::vector<int64_t> ids;
  while (completed < nreq)
  {
    bool progress = false;
    // Retire finished units (any order: units on different streams finish
    // independently).
    for (auto it = inflight.begin(); it != inflight.end();)
    {
      const hipError_t q = hipEventQuery(sv->slots[(size_t)it->slot].done);
      if (q == hipErrorNotReady)
      {
        ++it;
        continue;
      }
      if (q != hipSuccess)
      {
        drain(sv, inflight);
        return rc::hip_error(q, "hipEventQuery", __FILE__, __LINE__);
      }
      const double t = seconds_since(t0);
      for (int64_t i : it->reqs) res[i].done_s = t;
      last_done = std::max(last_done, t);
      completed += (int64_t)it->reqs.size();
      sv->arena.release(it->off, it->bytes);
      free_slots.push_back(it->slot);
      it = inflight.erase(it);
      progress = true;
    }
    // Admit waiting requests in arrival order.
    while (next < nreq && !free_slots.empty())
    {
      const double now = seconds_since(t0);
      if (cfg.honor_arrivals && req[next].arrival_s > now) break;
      ids.clear();
      size_t bytes = 0;
      const bool batch = is_batchable(cfg, req[next]);
      if (!batch)
      {
        bytes = single_ws(req[next]);
        if (bytes > sv->arena.size)
        {
          res[next].status = RAJA_CDNA4_ERROR_ADMISSION;
          res[next].start_s = res[next].done_s = now;
          ++next;
          ++completed;
          progress = true;
          continue;
        }
        ids.push_back(next);
      }
      else
      {
        // Waiting batchable requests, in order, while the batch's bound stays
        // within the arena and max_batch.
        size_t bound = kBatchFixedBound;
        for (int64_t j = next; j < nreq && (int64_t)ids.size() < cfg.max_batch &&
                               is_batchable(cfg, req[j]) &&
                               (!cfg.honor_arrivals || req[j].arrival_s <= now);
             ++j)
        {
          bound += batch_bytes_bound(req[j]);
          if (bound > sv->arena.size && !ids.empty()) break;
          ids.push_back(j);
        }
        sv->plan.build(req, ids);
        bytes = sv->plan.total;
        if (bytes > sv->arena.size)
        {
          res[next].status = RAJA_CDNA4_ERROR_ADMISSION;
          res[next].start_s = res[next].done_s = now;
          ++next;
          ++completed;
          progress = true;
          continue;
        }
      }
      size_t off = 0;
      if (!sv->arena.alloc(bytes, &off)) break;  // wait for memory to come back
      const int slot = free_slots.back();
      free_slots.pop_back();
      hipStream_t s = sv->streams[(size_t)(next_stream % nstreams)];
      next_stream = (next_stream + 1) % nstreams;
      char* dev = sv->arena.base + off;
      const int st = batch ? launch_batch(sv, sv->slots[(size_t)slot], dev, s)
                           : launch_single(req[next], dev, s);
      if (st != RAJA_CDNA4_SUCCESS)
      {
        sv->arena.release(off, bytes);
        drain(sv, inflight);
        return st;
      }
      const hipError_t e = hipEventRecord(sv->slots[(size_t)slot].done, s);
      if (e != hipSuccess)
      {
        sv->arena.release(off, bytes);
        drain(sv, inflight);
        return rc::hip_error(e, "hipEventRecord", __FILE__, __LINE__);
      }
      Unit u;
      u.slot = slot;
      u.id = units++;
      u.off = off;
      u.bytes = bytes;
      u.reqs = ids;
      for (int64_t i : ids)
      {
        res[i].unit = u.id;
        res[i].unit_size = (int)ids.size();
        res[i].start_s = now;
      }
      inflight.push_back(std::move(u));
      next += (int64_t)ids.size();
      progress = true;
    }
    if (!progress) std::this_thread::yield();
  }
  if (wall_s) *wall_s = last_done;
  return RAJA_CDNA4_SUCCESS;
}

}  // namespace

extern "C" {

int raja_cdna4_server_create(const raja_cdna4_serve_config* config, raja_cdna4_server_s** out)
{
  if (!config || !out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL argument");
  const raja_cdna4_serve_config& c = *config;
  if ((c.mode != RAJA_CDNA4_SERVE_BATCHED && c.mode != RAJA_CDNA4_SERVE_PER_REQUEST) ||
      c.streams < 1 || c.streams > 16 || c.max_inflight < 1 || c.max_inflight > 64 ||
      c.max_batch < 1 || c.batch_limit < 0 || c.arena_bytes < (int64_t)kAlign)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad server configuration");
  }
  if (c.batch_limit > rws::sum_batch_max())
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT,
                         "batch_limit %lld exceeds the batched sum's %lld elements",
                         (long long)c.batch_limit, (long long)rws::sum_batch_max());
  }
  raja_cdna4_server_s* sv = new raja_cdna4_server_s();
  sv->cfg = c;
  auto fail = [&](hipError_t e, const char* what) {
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    for (Slot& sl : sv->slots)
    {
      if (sl.done) (void)hipEventDestroy(sl.done);
    }
    if (sv->arena.base) (void)hipFree(sv->arena.base);
    delete sv;
    (void)hipGetLastError();
    return rc::hip_error(e, what, __FILE__, __LINE__);
  };
  hipError_t e = hipGetDevice(&sv->device);
  if (e != hipSuccess) return fail(e, "hipGetDevice");
  sv->arena.size = rc::align_up((size_t)c.arena_bytes, kAlign);
  e = hipMalloc((void**)&sv->arena.base, sv->arena.size);
  if (e != hipSuccess)
  {
    sv->arena.base = nullptr;
    fail(e, "hipMalloc(arena)");
    return rc::set_error(RAJA_CDNA4_ERROR_OUT_OF_MEMORY, "arena of %lld bytes not available",
                         (long long)c.arena_bytes);
  }
  const int ns = c.mode == RAJA_CDNA4_SERVE_BATCHED ? c.streams : 1;
  for (int k = 0; k < ns; ++k)
  {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(e, "hipStreamCreateWithFlags");
    sv->streams.push_back(s);
  }
  sv->slots.resize((size_t)c.max_inflight);
  for (Slot& sl : sv->slots)
  {
    e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e != hipSuccess) return fail(e, "hipEventCreateWithFlags");
  }
  *out = sv;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_server_run(raja_cdna4_server_s* sv, const raja_cdna4_request* requests, int64_t nreq,
                          raja_cdna4_result* results, double* wall_s)
{
  if (!sv || nreq < 0 || (nreq > 0 && (!requests || !results)))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad server run arguments");
  }
  if (wall_s) *wall_s = 0.0;
  return run(sv, requests, nreq, results, wall_s);
}

int raja_cdna4_server_destroy(raja_cdna4_server_s* sv)
{
  if (!sv) return RAJA_CDNA4_SUCCESS;
  for (hipStream_t s : sv->streams) RC_HIP(hipStreamSynchronize(s));
  for (hipStream_t s : sv->streams) RC_HIP(hipStreamDestroy(s));
  for (Slot& sl : sv->slots)
  {
    if (sl.done) RC_HIP(hipEventDestroy(sl.done));
    if (sl.staging) RC_HIP(hipHostFree(sl.staging));
  }
  if (sv->arena.base) RC_HIP(hipFree(sv->arena.base));
  delete sv;
  return RAJA_CDNA4_SUCCESS;
}

}  // extern "C"
