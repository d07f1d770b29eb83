// Request server: arena-backed admission, coalescing of small requests into
// batched launches, large requests alone.
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
//     (tile records carry each request's output offset), sorts of up to
//     8192 pairs with one block per request in LDS and longer ones with the
//     segmented onesweep (one launch per digit pass over all their tiles);
//   * units with scans or radix passes run on stream 0 (lookback launches
//     run one at a time on the device anyway), the others round-robin over
//     the remaining streams; a batch's tables upload on a stream of their
//     own, so an upload overlaps the units before it;
//   * completion is observed by polling one event per unit, so new arrivals
//     are admitted while earlier units run.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
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

// A unit slot: the pinned staging of a batch's tables, the events recorded
// after their upload and after the batch's block sorts, and the completion
// event.  A slot is reused only after its unit retired.
struct Slot
{
  hipEvent_t done = nullptr;
  hipEvent_t uploaded = nullptr;
  hipEvent_t sorted = nullptr;
  char* staging = nullptr;
  size_t staging_cap = 0;
};

// A request waiting for admission: its ticket (run(): its index) and the
// time it entered the queue.
struct Pending
{
  raja_cdna4_request req;
  int64_t ticket = 0;
  double arrival = 0.0;
};

struct Unit
{
  int slot = -1;
  int id = 0;
  size_t off = 0, bytes = 0;
  std::vector<int64_t> tickets;
  double start = 0.0;
};

bool is_batchable(const raja_cdna4_serve_config& cfg, const raja_cdna4_request& r)
{
  if (cfg.mode != RAJA_CDNA4_SERVE_BATCHED) return false;
  if (r.kind == RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32)
  {
    return r.n <= rws::block_sort_max() || r.n <= cfg.batch_limit;
  }
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

// Algorithmic bytes of a request: 8 per element for a sum or a scan, 68 per
// pair for a sort (four passes of 16 and the histogram read).
int64_t request_bytes(const raja_cdna4_request& r)
{
  return r.kind == RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32 ? 68 * r.n : 8 * r.n;
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
    default:
    {
      if (r.n <= rws::block_sort_max()) return sizeof(rs::SortSeg);
      // segmented onesweep: records, descriptors, a histogram row and the
      // alternate keys and values
      const int64_t t = rc::ceil_div(r.n, rsort::seg_tile());
      return (size_t)t * sizeof(rs::SortTile) + rsort::seg_desc_bytes(t) + rsort::seg_hist_bytes(1) +
             2 * rc::align_up((size_t)r.n * 4, kAlign);
    }
  }
}
constexpr size_t kBatchFixedBound = 6 * 16 + 8 * kAlign + 16;

// Host image of one batch: the records of each kind plus the layout of the
// device copy (records first, then partials and descriptors).
struct BatchPlan
{
  std::vector<rs::TileRec> sum_tiles, scan_tiles;
  std::vector<uint32_t> sum_first, sum_count;
  std::vector<int64_t> sum_out;
  std::vector<rs::SortSeg> sorts;       // one block each (block_sort_max() keys or fewer)
  std::vector<rs::SortTile> sort_tiles;  // tiles of the longer sorts (segmented onesweep)
  int64_t sort_ranges = 0, alt_elems = 0;
  uintptr_t sum_base = 0, scan_base = 0, sort_base = 0;
  size_t o_sum_tiles = 0, o_sum_first = 0, o_sum_count = 0, o_sum_out = 0, o_scan_tiles = 0,
         o_sorts = 0, o_sort_tiles = 0, rec_bytes = 0, o_partials = 0, o_desc = 0, o_sdesc = 0,
         o_hist = 0, o_alt_k = 0, o_alt_v = 0, total = 0;

  void clear()
  {
    sum_tiles.clear();
    scan_tiles.clear();
    sum_first.clear();
    sum_count.clear();
    sum_out.clear();
    sorts.clear();
    sort_tiles.clear();
    sort_ranges = alt_elems = 0;
  }

  static uintptr_t lowest(uintptr_t a, const void* p)
  {
    const uintptr_t v = (uintptr_t)p;
    return (a == 0 || v < a) ? v : a;
  }

  // Builds the records of the requests `reqs` (all batchable).
  void build(const std::vector<const raja_cdna4_request*>& reqs)
  {
    clear();
    sum_base = scan_base = sort_base = 0;
    for (const raja_cdna4_request* rp : reqs)
    {
      const raja_cdna4_request& r = *rp;
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
    for (const raja_cdna4_request* rp : reqs)
    {
      const raja_cdna4_request& r = *rp;
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
      else if (r.n && r.n <= rws::block_sort_max())
      {
        sorts.push_back(rs::SortSeg{(int64_t)(((uintptr_t)r.in - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.in_vals - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.out - sort_base) / 4),
                                    (int64_t)(((uintptr_t)r.out_vals - sort_base) / 4), r.n});
      }
      else if (r.n)
      {
        // offsets from sort_base (inputs, outputs) and from the batch's
        // alternate buffers (ak = av)
        const int64_t ik = (int64_t)(((uintptr_t)r.in - sort_base) / 4);
        const int64_t iv = (int64_t)(((uintptr_t)r.in_vals - sort_base) / 4);
        const int64_t ok = (int64_t)(((uintptr_t)r.out - sort_base) / 4);
        const int64_t ov = (int64_t)(((uintptr_t)r.out_vals - sort_base) / 4);
        const int64_t tile = rsort::seg_tile();
        for (int64_t j = 0; j * tile < r.n; ++j)
        {
          sort_tiles.push_back(rs::SortTile{ik, iv, ok, ov, alt_elems, alt_elems, r.n, (uint32_t)j,
                                            (uint32_t)sort_ranges});
        }
        ++sort_ranges;
        alt_elems += (r.n + 63) / 64 * 64;
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
    o_sort_tiles = place(sizeof(rs::SortTile) * sort_tiles.size());
    o_sum_out = place(sizeof(int64_t) * sum_out.size());
    o_sum_first = place(sizeof(uint32_t) * sum_first.size());
    o_sum_count = place(sizeof(uint32_t) * sum_count.size());
    rec_bytes = o;
    o = rc::align_up(o, kAlign);
    o_partials = o;
    o += rc::align_up(sizeof(double) * (sum_tiles.size() + 1), kAlign);
    o_desc = o;
    o += rws::exclusive_i32_ws_bytes((int64_t)scan_tiles.size(), true);
    o_sdesc = o;
    o += rc::align_up(rsort::seg_desc_bytes((int64_t)sort_tiles.size()), kAlign);
    o_hist = o;
    o += rc::align_up(rsort::seg_hist_bytes(sort_ranges), kAlign);
    o_alt_k = o;
    o += rc::align_up((size_t)alt_elems * 4, kAlign);
    o_alt_v = o;
    o += rc::align_up((size_t)alt_elems * 4, kAlign);
    total = o;
  }

  void write(char* dst) const
  {
    ::memcpy(dst + o_sum_tiles, sum_tiles.data(), sizeof(rs::TileRec) * sum_tiles.size());
    ::memcpy(dst + o_scan_tiles, scan_tiles.data(), sizeof(rs::TileRec) * scan_tiles.size());
    ::memcpy(dst + o_sorts, sorts.data(), sizeof(rs::SortSeg) * sorts.size());
    ::memcpy(dst + o_sort_tiles, sort_tiles.data(), sizeof(rs::SortTile) * sort_tiles.size());
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
  // Batched mode: the stream that uploads batch tables, so a unit's upload
  // runs while the units before it compute instead of between them on
  // their stream.
  hipStream_t upload = nullptr;
  std::vector<Slot> slots;
  BatchPlan plan;

  // Open serving (raja_cdna4_server_start / submit / poll / stop): the
  // scheduling thread takes submitted requests from `inbox` and posts
  // completions to `outbox`.
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<Pending> inbox;
  std::deque<raja_cdna4_completion> outbox;
  int64_t next_ticket = 0;
  int64_t outstanding = 0;  // submitted, not yet posted to outbox
  bool running = false, stopping = false;
  Clock::time_point t0;
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
    RC_HIP(hipMemcpyAsync(dev, slot.staging, p.rec_bytes, hipMemcpyHostToDevice, sv->upload));
    RC_HIP(hipEventRecord(slot.uploaded, sv->upload));
    RC_HIP(hipStreamWaitEvent(s, slot.uploaded, 0));
  }
  // The block sorts (a few dozen blocks, about 35 us) run on the upload
  // stream, after the upload, beside the unit's other kernels; the unit's
  // stream waits for them at its end.
  const bool side = !p.sorts.empty() && (!p.sum_out.empty() || !p.scan_tiles.empty() ||
                                         !p.sort_tiles.empty());
  if (!p.sorts.empty())
  {
    hipStream_t bs = side ? sv->upload : s;
    RC_TRY(rws::block_sort_u32_i32((const uint32_t*)p.sort_base, (const rs::SortSeg*)(dev + p.o_sorts),
                                   (int64_t)p.sorts.size(), bs));
    if (side) RC_HIP(hipEventRecord(slot.sorted, bs));
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
  if (!p.sort_tiles.empty())
  {
    void* base = (void*)p.sort_base;
    RC_TRY(rsort::seg_sort_table(RAJA_CDNA4_U32, 4, 0, base, base, base, base, dev + p.o_alt_k,
                                 dev + p.o_alt_v, (const rs::SortTile*)(dev + p.o_sort_tiles),
                                 (int64_t)p.sort_tiles.size(), p.sort_ranges,
                                 (uint32_t*)(dev + p.o_hist), (uint32_t*)(dev + p.o_sdesc), s));
  }
  if (side) RC_HIP(hipStreamWaitEvent(s, slot.sorted, 0));
  return RAJA_CDNA4_SUCCESS;
}

double seconds_since(Clock::time_point t0)
{
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Admission and retirement over a queue of ready requests, shared by the
// closed replay (run) and the open server's thread.  `finish(ticket,
// result)` reports each request once.
struct Scheduler
{
  raja_cdna4_server_s* sv;
  std::deque<Pending> ready;
  std::deque<Unit> inflight;
  std::vector<int> free_slots;
  int nstreams = 1, next_stream = 0, units = 0;
  Clock::time_point t0;

  Scheduler(raja_cdna4_server_s* s, Clock::time_point start) : sv(s), t0(start)
  {
    for (int k = (int)sv->slots.size() - 1; k >= 0; --k) free_slots.push_back(k);
    nstreams = sv->cfg.mode == RAJA_CDNA4_SERVE_BATCHED ? (int)sv->streams.size() : 1;
  }

  bool idle() const { return ready.empty() && inflight.empty(); }

  template<typename Finish>
  void fail_unit(const Unit& u, int status, Finish& finish)
  {
    const double t = seconds_since(t0);
    for (int64_t tk : u.tickets)
    {
      finish(tk, raja_cdna4_result{status, u.id, (int)u.tickets.size(), 0, u.start, t});
    }
  }

  // Retires finished units; returns whether any did, or an error status
  // (then every unit in flight has been waited for and failed).
  template<typename Finish>
  int retire(Finish& finish, bool* progress)
  {
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
        const int st = rc::hip_error(q, "hipEventQuery", __FILE__, __LINE__);
        for (hipStream_t s : sv->streams) (void)hipStreamSynchronize(s);
        if (sv->upload) (void)hipStreamSynchronize(sv->upload);
        for (Unit& u : inflight)
        {
          fail_unit(u, st, finish);
          sv->arena.release(u.off, u.bytes);
        }
        inflight.clear();
        return st;
      }
      const double t = seconds_since(t0);
      for (int64_t tk : it->tickets)
      {
        finish(tk, raja_cdna4_result{RAJA_CDNA4_SUCCESS, it->id, (int)it->tickets.size(), 0,
                                     it->start, t});
      }
      sv->arena.release(it->off, it->bytes);
      free_slots.push_back(it->slot);
      it = inflight.erase(it);
      *progress = true;
    }
    return RAJA_CDNA4_SUCCESS;
  }

  // Launches units from the front of `ready` while slots and arena memory
  // last: a request that is not batchable alone, else the longest run of
  // batchable requests within max_batch and the arena.
  template<typename Finish>
  int admit(Finish& finish, bool* progress)
  {
    const raja_cdna4_serve_config& cfg = sv->cfg;
    std::vector<const raja_cdna4_request*> reqs;
    while (!ready.empty() && !free_slots.empty())
    {
      const double now = seconds_since(t0);
      const Pending& front = ready.front();
      size_t take = 1, bytes = 0;
      const bool batch = is_batchable(cfg, front.req);
      if (!batch)
      {
        bytes = single_ws(front.req);
      }
      else
      {
        size_t bound = kBatchFixedBound;
        int64_t work = 0;
        reqs.clear();
        for (size_t j = 0; j < ready.size() && (int64_t)reqs.size() < cfg.max_batch &&
                           is_batchable(cfg, ready[j].req);
             ++j)
        {
          if (cfg.batch_bytes > 0 && work >= cfg.batch_bytes) break;
          bound += batch_bytes_bound(ready[j].req);
          if (bound > sv->arena.size && !reqs.empty()) break;
          reqs.push_back(&ready[j].req);
          work += request_bytes(ready[j].req);
        }
        take = reqs.size();
        sv->plan.build(reqs);
        bytes = sv->plan.total;
      }
      if (bytes > sv->arena.size)
      {
        finish(front.ticket, raja_cdna4_result{RAJA_CDNA4_ERROR_ADMISSION, -1, 0, 0, now, now});
        ready.pop_front();
        *progress = true;
        continue;
      }
      size_t off = 0;
      if (!sv->arena.alloc(bytes, &off)) break;  // wait for memory to come back
      const int slot = free_slots.back();
      free_slots.pop_back();
      // Units with lookback kernels (scans, radix passes) run one at a time on
      // the device anyway (rc::LookbackGate): they all go to stream 0, so the
      // gate never waits across streams; the others (sums, block sorts) go
      // round-robin over the remaining streams.
      const bool lookback =
          batch ? (!sv->plan.scan_tiles.empty() || !sv->plan.sort_tiles.empty())
                : (front.req.kind != RAJA_CDNA4_REQ_SUM_F64 && front.req.n > 1);
      hipStream_t s = sv->streams[0];
      if (!lookback && nstreams > 1)
      {
        s = sv->streams[(size_t)(1 + next_stream % (nstreams - 1))];
        next_stream = (next_stream + 1) % (nstreams - 1);
      }
      char* dev = sv->arena.base + off;
      Unit u;
      u.slot = slot;
      u.id = units++;
      u.off = off;
      u.bytes = bytes;
      u.start = now;
      for (size_t j = 0; j < take; ++j) u.tickets.push_back(ready[j].ticket);
      int st = batch ? launch_batch(sv, sv->slots[(size_t)slot], dev, s)
                     : launch_single(front.req, dev, s);
      if (st == RAJA_CDNA4_SUCCESS)
      {
        const hipError_t e = hipEventRecord(sv->slots[(size_t)slot].done, s);
        if (e != hipSuccess) st = rc::hip_error(e, "hipEventRecord", __FILE__, __LINE__);
      }
      for (size_t j = 0; j < take; ++j) ready.pop_front();
      *progress = true;
      if (st != RAJA_CDNA4_SUCCESS)
      {
        (void)hipStreamSynchronize(s);
        if (sv->upload) (void)hipStreamSynchronize(sv->upload);
        fail_unit(u, st, finish);
        sv->arena.release(off, bytes);
        free_slots.push_back(slot);
        return st;
      }
      inflight.push_back(std::move(u));
    }
    return RAJA_CDNA4_SUCCESS;
  }
};

// The closed replay: requests enter the ready queue at their arrival times
// (all at once without honor_arrivals) and the call returns when every one
// completed.
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
  Scheduler sch(sv, Clock::now());
  int64_t next = 0, completed = 0;
  double last_done = 0.0;
  int first_error = RAJA_CDNA4_SUCCESS;
  auto finish = [&](int64_t tk, const raja_cdna4_result& r) {
    res[tk] = r;
    last_done = std::max(last_done, r.done_s);
    ++completed;
  };
  while (completed < nreq)
  {
    bool progress = false;
    const double now = seconds_since(sch.t0);
    while (next < nreq && (!cfg.honor_arrivals || req[next].arrival_s <= now))
    {
      sch.ready.push_back(Pending{req[next], next, req[next].arrival_s});
      ++next;
    }
    int st = sch.retire(finish, &progress);
    if (st == RAJA_CDNA4_SUCCESS) st = sch.admit(finish, &progress);
    if (st != RAJA_CDNA4_SUCCESS)
    {
      // wait for what is in flight and report the first failure
      first_error = st;
      for (hipStream_t s : sv->streams) (void)hipStreamSynchronize(s);
      bool p2 = false;
      (void)sch.retire(finish, &p2);
      break;
    }
    if (!progress) std::this_thread::yield();
  }
  if (wall_s) *wall_s = last_done;
  return first_error;
}

// The open server's scheduling thread.
void serve_loop(raja_cdna4_server_s* sv)
{
  (void)hipSetDevice(sv->device);
  Scheduler sch(sv, sv->t0);
  auto finish = [&](int64_t tk, const raja_cdna4_result& r) {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->outbox.push_back(raja_cdna4_completion{tk, r});
    --sv->outstanding;
    sv->cv_out.notify_all();
  };
  while (true)
  {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (sch.idle() && sv->inbox.empty())
      {
        if (sv->stopping) break;
        sv->cv_in.wait_for(lk, std::chrono::milliseconds(1));
      }
      while (!sv->inbox.empty())
      {
        sch.ready.push_back(sv->inbox.front());
        sv->inbox.pop_front();
      }
    }
    bool progress = false;
    (void)sch.retire(finish, &progress);  // failures are reported per request
    (void)sch.admit(finish, &progress);
    if (!progress && !sch.idle()) std::this_thread::yield();
  }
}

}  // namespace

extern "C" {

int raja_cdna4_server_create(const raja_cdna4_serve_config* config, raja_cdna4_server_s** out)
{
  if (!config || !out) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL argument");
  const raja_cdna4_serve_config& c = *config;
  if ((c.mode != RAJA_CDNA4_SERVE_BATCHED && c.mode != RAJA_CDNA4_SERVE_PER_REQUEST) ||
      c.streams < 1 || c.streams > 16 || c.max_inflight < 1 || c.max_inflight > 64 ||
      c.max_batch < 1 || c.batch_limit < 0 || c.arena_bytes < (int64_t)kAlign || c.batch_bytes < 0)
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
    if (sv->upload) (void)hipStreamDestroy(sv->upload);
    for (Slot& sl : sv->slots)
    {
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.uploaded) (void)hipEventDestroy(sl.uploaded);
      if (sl.sorted) (void)hipEventDestroy(sl.sorted);
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
  if (c.mode == RAJA_CDNA4_SERVE_BATCHED)
  {
    e = hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(e, "hipStreamCreateWithFlags");
  }
  sv->slots.resize((size_t)c.max_inflight);
  for (Slot& sl : sv->slots)
  {
    e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e != hipSuccess) return fail(e, "hipEventCreateWithFlags");
    e = hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming);
    if (e != hipSuccess) return fail(e, "hipEventCreateWithFlags");
    e = hipEventCreateWithFlags(&sl.sorted, hipEventDisableTiming);
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
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->running)
    {
      return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "server is serving submitted requests");
    }
  }
  return run(sv, requests, nreq, results, wall_s);
}

int raja_cdna4_server_start(raja_cdna4_server_s* sv)
{
  if (!sv) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL server");
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->running) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "server already started");
  sv->arena.reset();
  sv->inbox.clear();
  sv->outbox.clear();
  sv->outstanding = 0;
  sv->stopping = false;
  sv->t0 = Clock::now();
  sv->running = true;
  sv->worker = std::thread(serve_loop, sv);
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_server_submit(raja_cdna4_server_s* sv, const raja_cdna4_request* reqs, int64_t n,
                             int64_t* tickets)
{
  if (!sv || n < 0 || (n > 0 && !reqs))
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad submit arguments");
  }
  for (int64_t i = 0; i < n; ++i) RC_TRY(check_request(reqs[i]));
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (!sv->running || sv->stopping)
    {
      return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "server not started");
    }
    const double now = seconds_since(sv->t0);
    for (int64_t i = 0; i < n; ++i)
    {
      Pending p;
      p.req = reqs[i];
      p.req.arrival_s = now;
      p.ticket = sv->next_ticket++;
      p.arrival = now;
      if (tickets) tickets[i] = p.ticket;
      sv->inbox.push_back(p);
    }
    sv->outstanding += n;
  }
  sv->cv_in.notify_one();
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_server_poll(raja_cdna4_server_s* sv, raja_cdna4_completion* out, int64_t max,
                           double timeout_s, int64_t* count)
{
  if (!sv || max < 0 || (max > 0 && !out) || !count)
  {
    return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "bad poll arguments");
  }
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->outbox.empty() && timeout_s > 0.0 && sv->outstanding > 0)
  {
    sv->cv_out.wait_for(lk, std::chrono::duration<double>(timeout_s),
                        [&] { return !sv->outbox.empty() || sv->outstanding == 0; });
  }
  int64_t k = 0;
  while (k < max && !sv->outbox.empty())
  {
    out[k++] = sv->outbox.front();
    sv->outbox.pop_front();
  }
  *count = k;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_server_stop(raja_cdna4_server_s* sv)
{
  if (!sv) return rc::set_error(RAJA_CDNA4_ERROR_INVALID_ARGUMENT, "NULL server");
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (!sv->running) return RAJA_CDNA4_SUCCESS;
    sv->stopping = true;
  }
  sv->cv_in.notify_one();
  sv->worker.join();  // the thread ends once every submitted request completed
  std::lock_guard<std::mutex> g(sv->mu);
  sv->running = false;
  return RAJA_CDNA4_SUCCESS;
}

int raja_cdna4_server_destroy(raja_cdna4_server_s* sv)
{
  if (!sv) return RAJA_CDNA4_SUCCESS;
  RC_TRY(raja_cdna4_server_stop(sv));
  for (hipStream_t s : sv->streams) RC_HIP(hipStreamSynchronize(s));
  if (sv->upload) RC_HIP(hipStreamSynchronize(sv->upload));
  for (hipStream_t s : sv->streams) RC_HIP(hipStreamDestroy(s));
  if (sv->upload) RC_HIP(hipStreamDestroy(sv->upload));
  for (Slot& sl : sv->slots)
  {
    if (sl.done) RC_HIP(hipEventDestroy(sl.done));
    if (sl.uploaded) RC_HIP(hipEventDestroy(sl.uploaded));
    if (sl.sorted) RC_HIP(hipEventDestroy(sl.sorted));
    if (sl.staging) RC_HIP(hipHostFree(sl.staging));
  }
  if (sv->arena.base) RC_HIP(hipFree(sv->arena.base));
  delete sv;
  return RAJA_CDNA4_SUCCESS;
}

}  // extern "C"
