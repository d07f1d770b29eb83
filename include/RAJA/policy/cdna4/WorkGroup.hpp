// cdna4 WorkGroup runners.
//
// Reference: include/RAJA/policy/hip/WorkGroup/WorkRunner.hpp:38-425 and
// Dispatcher.hpp:42-115.  The reference's unordered HIP runner launches one
// kernel with blockIdx.y selecting a loop and x sized by the *average* loop
// length, so long loops are under-parallelised and short ones leave blocks
// idle.  Here the loops are cut into fixed tiles of 8*B iterations listed in
// a device tile table, built once per WorkGroup; one kernel runs one block
// per tile, so 131072-iteration faces and 8-iteration corners share the GPU
// evenly.  Ordered runs launch one kernel per loop in (reverse) enqueue
// order, as the reference does.
//
// Dispatch: direct_dispatch<camp::list<Seg, Body>...> switches on the body's
// position in the list at compile time (bodies inlined); the indirect
// dispatch policies call through a device function pointer fetched once per
// body type.
#pragma once

#include <cstring>
#include <vector>

#include "RAJA/pattern/WorkGroup.hpp"
#include "RAJA/policy/cdna4/forall.hpp"

namespace RAJA
{
namespace policy
{
namespace cdna4
{

template<int BLOCK_SIZE, bool ASYNC = false>
struct cdna4_work_t
{
  static_assert(BLOCK_SIZE % 64 == 0 && BLOCK_SIZE <= 1024, "block size: multiple of 64, <= 1024");
  using resource_type = resources::Cdna4;
  static constexpr int block_size = BLOCK_SIZE;
  static constexpr bool async = ASYNC;
};

// Unordered runner over the tile table.
struct unordered_cdna4_loop_tiles
{
};

}  // namespace cdna4
}  // namespace policy

template<int B>
using cdna4_work = policy::cdna4::cdna4_work_t<B, false>;
template<int B>
using cdna4_work_async = policy::cdna4::cdna4_work_t<B, true>;
using policy::cdna4::unordered_cdna4_loop_tiles;
// Name of the reference's unordered HIP runner, mapped to the tile runner.
using unordered_cdna4_loop_y_block_iter_x_threadblock_average = unordered_cdna4_loop_tiles;

namespace detail
{
namespace workgroup
{

template<typename Body, typename Index, typename... Args>
__device__ void device_call(const void* body, Index i, Args... args)
{
  (*static_cast<const Body*>(body))(i, args...);
}

template<typename Index, typename... Args>
using device_call_t = void (*)(const void*, Index, Args...);

template<typename Body, typename Index, typename... Args>
__global__ void device_fn_kernel(device_call_t<Index, Args...>* out)
{
  *out = &device_call<Body, Index, Args...>;
}

template<typename Body, typename Index, typename... Args>
void* device_fn_of()
{
  static void* cached = nullptr;
  if (!cached)
  {
    device_call_t<Index, Args...>* d = nullptr;
    resources::detail::check_lib(raja_cdna4_device_malloc((void**)&d, sizeof(*d)), "fn slot");
    hipLaunchKernelGGL((device_fn_kernel<Body, Index, Args...>), dim3(1), dim3(1), 0, 0, d);
    resources::detail::check_hip(hipGetLastError(), "device_fn_kernel");
    device_call_t<Index, Args...> h = nullptr;
    resources::detail::check_hip(hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost), "fn copy");
    resources::detail::check_lib(raja_cdna4_device_free(d), "fn slot");
    cached = (void*)h;
  }
  return cached;
}

// ---- device dispatch -----------------------------------------------------------

template<typename List, typename Index, typename... Args>
struct DirectCall;
template<typename Index, typename... Args>
struct DirectCall<camp::list<>, Index, Args...>
{
  __device__ static void call(uint32_t, const void*, Index, Args...) {}
};
template<typename Seg, typename Body, typename... Rest, typename Index, typename... Args>
struct DirectCall<camp::list<camp::list<Seg, Body>, Rest...>, Index, Args...>
{
  __device__ static void call(uint32_t k, const void* body, Index i, Args... args)
  {
    if (k == 0)
    {
      (*static_cast<const Body*>(body))(i, args...);
    }
    else
    {
      DirectCall<camp::list<Rest...>, Index, Args...>::call(k - 1, body, i, args...);
    }
  }
};

template<typename Dispatch, typename Index, typename... Args>
struct Dispatcher
{
  __device__ static void call(const LoopRecord& L, const void* body, Index i, Args... args)
  {
    reinterpret_cast<device_call_t<Index, Args...>>(L.device_fn)(body, i, args...);
  }
};
template<typename... RC, typename Index, typename... Args>
struct Dispatcher<direct_dispatch<RC...>, Index, Args...>
{
  __device__ static void call(const LoopRecord& L, const void* body, Index i, Args... args)
  {
    DirectCall<camp::list<RC...>, Index, Args...>::call(L.type_index, body, i, args...);
  }
};

struct Tile
{
  uint32_t loop;
  uint32_t pad;
  int64_t start;
  int64_t end;
};

// Iterations per thread of one tile (a tile is kTileItems * B iterations).
// Halo pack + unpack at 256^3 with B = 256: 0.206 ms at 2, 0.169 at 4,
// 0.161 at 8.
constexpr int kTileItems = 8;

template<int B, typename Dispatch, typename Index, typename... Args>
__global__ __launch_bounds__(B) void tiles_kernel(const Tile* __restrict__ tiles,
                                                  const LoopRecord* __restrict__ loops,
                                                  const unsigned char* __restrict__ bytes,
                                                  Args... args)
{
  const Tile t = tiles[blockIdx.x];
  const LoopRecord L = loops[t.loop];
  const void* body = bytes + L.body_offset;
  if (t.end - t.start == (int64_t)kTileItems * B)
  {
    // A full tile: the thread's iterations in straight-line code, so the
    // loads of later iterations can issue before earlier ones complete
    // (a gather body is otherwise one dependent load chain per iteration).
#pragma unroll
    for (int u = 0; u < kTileItems; ++u)
    {
      Dispatcher<Dispatch, Index, Args...>::call(L, body, Index(L.begin + t.start + u * B + threadIdx.x),
                                                 args...);
    }
  }
  else
  {
    for (int64_t k = t.start + threadIdx.x; k < t.end; k += B)
    {
      Dispatcher<Dispatch, Index, Args...>::call(L, body, Index(L.begin + k), args...);
    }
  }
}

template<int B, typename Dispatch, typename Index, typename... Args>
__global__ __launch_bounds__(B) void loop_kernel(const LoopRecord* __restrict__ loops,
                                                 uint32_t l,
                                                 const unsigned char* __restrict__ bytes,
                                                 Args... args)
{
  const LoopRecord L = loops[l];
  const void* body = bytes + L.body_offset;
  const int64_t stride = (int64_t)gridDim.x * B;
  for (int64_t k = (int64_t)blockIdx.x * B + threadIdx.x; k < L.len; k += stride)
  {
    Dispatcher<Dispatch, Index, Args...>::call(L, body, Index(L.begin + k), args...);
  }
}

// Uploaded tables of recently released groups (device block and its pinned
// host image), per block size and thread, most recent last.  A group whose
// host image matches one of them byte for byte (same bodies, loop records
// and tile table, as when a time step re-enqueues the same loops) takes it
// over without an upload.  Entries are idle: their stream was synchronised
// when they were released.
struct CachedTables
{
  unsigned char* d_all = nullptr;
  unsigned char* h_all = nullptr;
  size_t total = 0;
};
constexpr size_t kTableCacheEntries = 4;
template<int B>
std::vector<CachedTables>& table_cache()
{
  // intentionally never destroyed: entries hold pool memory until exit
  thread_local std::vector<CachedTables>* c = new std::vector<CachedTables>();
  return *c;
}

// Device copies of a group's loop bodies and records and, for unordered
// runs, its tile table: built at the group's first run, kept for later
// runs, released when the loop set changes or the group goes away.  The
// three parts are staged in one pinned host buffer and uploaded with one
// asynchronous copy on the run's stream (no host wait: the staging buffer
// lives until the tables are released, after the stream is synchronised).
template<int B>
struct DeviceTables
{
  unsigned char* d_all = nullptr;  // bodies, records, tiles in one device block
  unsigned char* h_all = nullptr;  // pinned staging of the same bytes
  unsigned char* d_bytes = nullptr;
  LoopRecord* d_loops = nullptr;
  Tile* d_tiles = nullptr;
  int64_t ntiles = 0;
  bool built = false;
  hipStream_t stream = nullptr;

  DeviceTables() = default;
  DeviceTables(const DeviceTables&) = delete;
  DeviceTables& operator=(const DeviceTables&) = delete;
  DeviceTables(DeviceTables&& o) noexcept { *this = std::move(o); }
  DeviceTables& operator=(DeviceTables&& o) noexcept
  {
    std::swap(d_all, o.d_all);
    std::swap(h_all, o.h_all);
    std::swap(d_bytes, o.d_bytes);
    std::swap(d_loops, o.d_loops);
    std::swap(d_tiles, o.d_tiles);
    std::swap(ntiles, o.ntiles);
    std::swap(built, o.built);
    std::swap(stream, o.stream);
    std::swap(total, o.total);
    return *this;
  }
  ~DeviceTables() { release(); }

  size_t total = 0;

  void release()
  {
    if (!built) return;
    (void)hipStreamSynchronize(stream);
    std::vector<CachedTables>& cache = table_cache<B>();
    cache.push_back(CachedTables{d_all, h_all, total});
    if (cache.size() > kTableCacheEntries)
    {
      (void)raja_cdna4_device_free(cache.front().d_all);
      (void)raja_cdna4_pinned_free(cache.front().h_all);
      cache.erase(cache.begin());
    }
    d_all = h_all = d_bytes = nullptr;
    total = 0;
    d_loops = nullptr;
    d_tiles = nullptr;
    ntiles = 0;
    built = false;
  }

  static size_t round16(size_t v) { return (v + 15) / 16 * 16; }

  // Upload the loop bodies, records and (for unordered runs) the tile table.
  template<typename Storage>
  void build(const Storage& s, resources::Cdna4 r, bool tiles, bool indirect)
  {
    stream = r.get_stream();
    std::vector<LoopRecord> recs(s.loops);
    if (indirect)
    {
      for (size_t l = 0; l < recs.size(); ++l) recs[l].device_fn = s.host_fns[l].device_fn();
    }
    std::vector<Tile> t;
    if (tiles)
    {
      constexpr int64_t TILE = (int64_t)kTileItems * B;
      for (size_t l = 0; l < recs.size(); ++l)
      {
        for (int64_t s0 = 0; s0 < recs[l].len; s0 += TILE)
        {
          t.push_back(Tile{(uint32_t)l, 0u, s0, std::min(recs[l].len, s0 + TILE)});
        }
      }
    }
    const size_t o_loops = round16(s.used + 16);
    const size_t o_tiles = o_loops + round16(sizeof(LoopRecord) * recs.size());
    total = o_tiles + sizeof(Tile) * t.size();
    std::vector<unsigned char> image(total, 0);
    std::memcpy(image.data(), s.bytes.get(), s.used);
    std::memcpy(image.data() + o_loops, recs.data(), sizeof(LoopRecord) * recs.size());
    if (!t.empty()) std::memcpy(image.data() + o_tiles, t.data(), sizeof(Tile) * t.size());
    std::vector<CachedTables>& cache = table_cache<B>();
    bool reused = false;
    for (size_t k = cache.size(); k-- > 0;)
    {
      if (cache[k].total == total && std::memcmp(cache[k].h_all, image.data(), total) == 0)
      {
        d_all = cache[k].d_all;  // the device block already holds this image
        h_all = cache[k].h_all;
        cache.erase(cache.begin() + (std::ptrdiff_t)k);
        reused = true;
        break;
      }
    }
    if (!reused)
    {
      resources::detail::check_lib(raja_cdna4_device_malloc((void**)&d_all, total), "wg tables");
      resources::detail::check_lib(raja_cdna4_pinned_malloc((void**)&h_all, total), "wg staging");
      std::memcpy(h_all, image.data(), total);
      resources::detail::check_hip(hipMemcpyAsync(d_all, h_all, total, hipMemcpyHostToDevice, stream),
                                   "wg tables upload");
    }
    d_bytes = d_all;
    d_loops = reinterpret_cast<LoopRecord*>(d_all + o_loops);
    d_tiles = t.empty() ? nullptr : reinterpret_cast<Tile*>(d_all + o_tiles);
    ntiles = (int64_t)t.size();
    built = true;
  }
};

}  // namespace workgroup

template<int B, bool A, typename Order, typename Dispatch, typename Alloc, typename Index,
         typename... Args>
struct WorkRunner<policy::cdna4::cdna4_work_t<B, A>, Order, Dispatch, Alloc, Index, Args...>
{
  using resource_type = resources::Cdna4;
  using dispatcher_type = workgroup::Dispatcher<Dispatch, Index, Args...>;
  // Nothing is allocated per run: the device tables belong to the group.
  using per_run_storage = int;
  static constexpr bool unordered = !std::is_same<Order, ordered>::value &&
                                    !std::is_same<Order, reverse_ordered>::value;
  static constexpr bool indirect =
      std::is_same<typename workgroup::dispatch_list<Dispatch>::type, void>::value;

  WorkRunner() = default;
  WorkRunner(WorkRunner&&) = default;
  WorkRunner& operator=(WorkRunner&&) = default;

  // Bodies are copied bytewise to the device at the first run.
  template<typename WorkContainer, typename Iterable, typename LoopBody>
  void enqueue(WorkContainer& storage, Iterable&& iter, LoopBody&& loop_body)
  {
    m_tables.release();  // the loop set changes: tables are rebuilt at the next run
    workgroup::enqueue_loop<true, Dispatch>(storage, std::forward<Iterable>(iter),
                                            std::forward<LoopBody>(loop_body));
  }

  template<typename WorkContainer>
  per_run_storage run(WorkContainer const& s, resource_type r, Args... args) const
  {
    if (s.size() == 0) return 0;
    hipStream_t st = r.get_stream();
    if (!m_tables.built)
    {
      m_tables.build(s, r, unordered, indirect);
    }
    else if (m_tables.stream != st)
    {
      // the tables were uploaded (and last read) on another stream
      resources::detail::check_hip(hipStreamSynchronize(m_tables.stream), "wg stream switch");
      m_tables.stream = st;
    }
    if (unordered)
    {
      if (m_tables.ntiles > 0)
      {
        hipLaunchKernelGGL((workgroup::tiles_kernel<B, Dispatch, Index, Args...>),
                           dim3((unsigned)m_tables.ntiles), dim3(B), 0, st, m_tables.d_tiles,
                           m_tables.d_loops, m_tables.d_bytes, args...);
        resources::detail::check_hip(hipGetLastError(), "workgroup tiles launch");
      }
    }
    else
    {
      static int cap = 0;
      if (cap == 0)
      {
        cap = RAJA::cdna4::detail::occupancy_grid(workgroup::loop_kernel<B, Dispatch, Index, Args...>, B);
      }
      const size_t n = s.size();
      for (size_t k = 0; k < n; ++k)
      {
        const size_t l = std::is_same<Order, reverse_ordered>::value ? n - 1 - k : k;
        int64_t grid = (s.loops[l].len + B - 1) / B;
        if (grid > cap) grid = cap;
        hipLaunchKernelGGL((workgroup::loop_kernel<B, Dispatch, Index, Args...>), dim3((unsigned)grid),
                           dim3(B), 0, st, m_tables.d_loops, (uint32_t)l, m_tables.d_bytes, args...);
      }
      resources::detail::check_hip(hipGetLastError(), "workgroup loop launch");
    }
    if (!A) r.wait();
    return 0;
  }

  void clear() { m_tables.release(); }

private:
  mutable workgroup::DeviceTables<B> m_tables;
};

}  // namespace detail
}  // namespace RAJA
