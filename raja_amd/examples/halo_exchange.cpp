// Halo exchange on the cdna4 policies: the workload of the reference's
// examples/tut_halo-exchange.cpp (its HIP sections, :1140-1520) -- a grid
// with a halo of width h on every side, num_vars FP64 variables set to
// var[i] = i + v each cycle, then 26 neighbour pack loops (buffer[i] =
// var[list[i]]) and 26 unpack loops (var[list[i]] = buffer[i]) per variable.
//
// Three ways of running the pack and unpack phases are timed:
//   forall     one RAJA::forall<cdna4_exec_async> per (neighbour, variable)
//              loop, a device synchronisation after each neighbour's packs
//              (the reference's forall section, :1255-1290);
//   ordered    a WorkGroup with RAJA::ordered: one launch per loop, one
//              synchronisation per phase;
//   unordered  a WorkGroup with unordered_cdna4_loop_tiles: one launch per
//              phase over a table of fixed 4*B-iteration tiles (the
//              reference's unordered_hip_loop_y_block_iter_x_threadblock_average,
//              :1380-1480, sizes its grid by the average loop length).
// Each variant checks every cell of every variable against the expected
// values afterwards.
//
// Usage: halo_exchange [gx gy gz halo num_vars num_cycles]   (default 256 256 256 2 10 5)
// Prints one JSON line.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "RAJA/RAJA.hpp"

namespace
{

constexpr int kNeighbors = 26;

// Neighbour directions in the reference's list order: 6 faces (-x,+x,-y,+y,
// -z,+z), 12 edges (xy, xz, yz pairs, minus before plus), 8 corners.
const int kDir[kNeighbors][3] = {
    {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},   {0, 0, -1},   {0, 0, 1},   {-1, -1, 0},
    {-1, 1, 0},  {1, -1, 0},  {1, 1, 0},   {-1, 0, -1}, {-1, 0, 1},   {1, 0, -1},  {1, 0, 1},
    {0, -1, -1}, {0, -1, 1},  {0, 1, -1},  {0, 1, 1},   {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1},
    {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};

// Index range along one axis of the cells a neighbour list covers.  Packing
// reads the interior layer next to the neighbour; unpacking writes the halo
// layer on that side.  d = 0 spans the whole interior.
void axis_range(int d, int g, int h, bool unpack, int* lo, int* hi)
{
  if (d == 0)
  {
    *lo = h;
    *hi = g + h;
  }
  else if (d < 0)
  {
    *lo = unpack ? 0 : h;
    *hi = unpack ? h : 2 * h;
  }
  else
  {
    *lo = unpack ? g + h : g;
    *hi = unpack ? g + 2 * h : g + h;
  }
}

// All 26 lists concatenated; lens[l] entries each, x fastest.
std::vector<int> make_lists(const int* g, int h, bool unpack, std::vector<int>& lens)
{
  const int sj = g[0] + 2 * h;
  const int sk = sj * (g[1] + 2 * h);
  std::vector<int> out;
  lens.assign(kNeighbors, 0);
  for (int l = 0; l < kNeighbors; ++l)
  {
    int r[3][2];
    for (int a = 0; a < 3; ++a) axis_range(kDir[l][a], g[a], h, unpack, &r[a][0], &r[a][1]);
    for (int k = r[2][0]; k < r[2][1]; ++k)
      for (int j = r[1][0]; j < r[1][1]; ++j)
        for (int i = r[0][0]; i < r[0][1]; ++i) out.push_back(i + j * sj + k * sk);
    lens[l] = (r[0][1] - r[0][0]) * (r[1][1] - r[1][0]) * (r[2][1] - r[2][0]);
  }
  return out;
}

uint64_t list_hash(const std::vector<int>& v)
{
  uint64_t hsh = 0;
  for (size_t i = 0; i < v.size(); ++i) hsh += (uint64_t)(uint32_t)v[i] * (uint64_t)(i + 1);
  return hsh;
}

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
  {
    std::fprintf(stderr, "halo_exchange: %s: %s\n", what, hipGetErrorString(e));
    std::exit(2);
  }
}

struct PackBody
{
  double* buf;
  const double* var;
  const int* list;
  RAJA_HOST_DEVICE void operator()(int i) const { buf[i] = var[list[i]]; }
};

struct UnpackBody
{
  const double* buf;
  double* var;
  const int* list;
  RAJA_HOST_DEVICE void operator()(int i) const { var[list[i]] = buf[i]; }
};

struct Timing
{
  std::vector<double> cycle_s, pack_ms, unpack_ms, pack_host_ms, unpack_host_ms;
  bool pass = false;
};

double median(std::vector<double> v)
{
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.0 : v[v.size() / 2];
}

}  // namespace

int main(int argc, char** argv)
{
  int g[3] = {256, 256, 256};
  int h = 2, nvars = 10, ncycles = 5;
  if (argc == 7)
  {
    g[0] = std::atoi(argv[1]);
    g[1] = std::atoi(argv[2]);
    g[2] = std::atoi(argv[3]);
    h = std::atoi(argv[4]);
    nvars = std::atoi(argv[5]);
    ncycles = std::atoi(argv[6]);
  }
  else if (argc != 1)
  {
    std::fprintf(stderr, "usage: %s [gx gy gz halo num_vars num_cycles]\n", argv[0]);
    return 1;
  }
  if (g[0] < h || g[1] < h || g[2] < h || h < 1 || nvars < 1 || ncycles < 1)
  {
    std::fprintf(stderr, "halo_exchange: grid dimensions must be >= halo width >= 1\n");
    return 1;
  }
  const int var_size = (g[0] + 2 * h) * (g[1] + 2 * h) * (g[2] + 2 * h);
  std::vector<int> plens, ulens;
  const std::vector<int> pl = make_lists(g, h, false, plens);
  const std::vector<int> ul = make_lists(g, h, true, ulens);
  const int64_t total = (int64_t)pl.size();

  // Expected contents after a cycle: var[i] = i + v, halo cells from the
  // matching interior cells.
  std::vector<int> src(var_size);
  for (int i = 0; i < var_size; ++i) src[i] = i;
  for (int64_t q = 0; q < total; ++q) src[ul[q]] = pl[q];

  auto res = RAJA::resources::Cdna4::get_default();
  hipStream_t st = res.get_stream();
  std::vector<double*> vars(nvars);
  for (auto& p : vars) check(hipMalloc(&p, sizeof(double) * var_size), "hipMalloc var");
  int *d_pl = nullptr, *d_ul = nullptr;
  double* d_buf = nullptr;
  check(hipMalloc(&d_pl, sizeof(int) * total), "hipMalloc lists");
  check(hipMalloc(&d_ul, sizeof(int) * total), "hipMalloc lists");
  check(hipMalloc(&d_buf, sizeof(double) * total * nvars), "hipMalloc buffers");
  check(hipMemcpy(d_pl, pl.data(), sizeof(int) * total, hipMemcpyHostToDevice), "copy lists");
  check(hipMemcpy(d_ul, ul.data(), sizeof(int) * total, hipMemcpyHostToDevice), "copy lists");
  std::vector<int64_t> loff(kNeighbors + 1, 0);
  for (int l = 0; l < kNeighbors; ++l) loff[l + 1] = loff[l] + plens[l];

  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");

  using EXEC = RAJA::cdna4_exec_async<256>;
  using SEG = RAJA::TypedRangeSegment<int>;
  using Disp = RAJA::direct_dispatch<camp::list<SEG, PackBody>, camp::list<SEG, UnpackBody>>;
  using PolU = RAJA::WorkGroupPolicy<RAJA::cdna4_work_async<256>, RAJA::unordered_cdna4_loop_tiles,
                                     RAJA::constant_stride_array_of_objects, Disp>;
  using PolO = RAJA::WorkGroupPolicy<RAJA::cdna4_work_async<256>, RAJA::ordered,
                                     RAJA::constant_stride_array_of_objects, Disp>;
  using PoolU = RAJA::WorkPool<PolU, int, RAJA::xargs<>, std::allocator<char>>;
  using PoolO = RAJA::WorkPool<PolO, int, RAJA::xargs<>, std::allocator<char>>;

  auto set_vars = [&]() {
    for (int v = 0; v < nvars; ++v)
    {
      double* var = vars[v];
      RAJA::forall<EXEC>(res, SEG(0, var_size), [=] RAJA_DEVICE(int i) { var[i] = i + v; });
    }
  };
  // buffer of neighbour l, variable v
  auto buf_of = [&](int l, int v) { return d_buf + loff[l] * nvars + (int64_t)v * plens[l]; };

  auto verify = [&]() {
    std::vector<double> hv(var_size);
    bool ok = true;
    for (int v = 0; v < nvars && ok; ++v)
    {
      check(hipMemcpy(hv.data(), vars[v], sizeof(double) * var_size, hipMemcpyDeviceToHost), "copy back");
      for (int i = 0; i < var_size; ++i)
      {
        if (hv[i] != (double)src[i] + v)
        {
          ok = false;
          break;
        }
      }
    }
    return ok;
  };

  // One cycle: set the variables, then the pack phase and the unpack phase.
  // Per phase: host_ms from its first enqueue to the end of its
  // synchronisation (what the caller waits), device_ms between events
  // recorded around its launches.
  auto run_variant = [&](int which) {
    Timing t;
    PoolU pool_u;
    PoolO pool_o;
    for (int c = 0; c < ncycles + 1; ++c)  // cycle 0 warms up
    {
      const auto c0 = std::chrono::steady_clock::now();
      set_vars();
      check(hipStreamSynchronize(st), "sync");
      double host_ms[2], dev_ms[2];
      for (int ph = 0; ph < 2; ++ph)
      {
        const bool unpack = ph == 1;
        const auto h0 = std::chrono::steady_clock::now();
        if (which == 0)
        {
          check(hipEventRecord(e0, st), "record");
          for (int l = 0; l < kNeighbors; ++l)
          {
            for (int v = 0; v < nvars; ++v)
            {
              if (unpack)
                RAJA::forall<EXEC>(res, SEG(0, plens[l]), UnpackBody{buf_of(l, v), vars[v], d_ul + loff[l]});
              else
                RAJA::forall<EXEC>(res, SEG(0, plens[l]), PackBody{buf_of(l, v), vars[v], d_pl + loff[l]});
            }
            if (!unpack) check(hipStreamSynchronize(st), "sync");  // one message per neighbour
          }
          check(hipEventRecord(e1, st), "record");
        }
        else
        {
          auto phase = [&](auto& pool) {
            for (int l = 0; l < kNeighbors; ++l)
            {
              for (int v = 0; v < nvars; ++v)
              {
                if (unpack)
                  pool.enqueue(SEG(0, plens[l]), UnpackBody{buf_of(l, v), vars[v], d_ul + loff[l]});
                else
                  pool.enqueue(SEG(0, plens[l]), PackBody{buf_of(l, v), vars[v], d_pl + loff[l]});
              }
            }
            auto group = pool.instantiate();
            check(hipEventRecord(e0, st), "record");
            auto site = group.run(res);
            check(hipEventRecord(e1, st), "record");
            check(hipStreamSynchronize(st), "sync");
          };
          if (which == 1)
            phase(pool_o);
          else
            phase(pool_u);
        }
        check(hipStreamSynchronize(st), "sync");
        host_ms[ph] = 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        dev_ms[ph] = ms;
      }
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
      if (c > 0)
      {
        t.cycle_s.push_back(secs);
        t.pack_ms.push_back(dev_ms[0]);
        t.unpack_ms.push_back(dev_ms[1]);
        t.pack_host_ms.push_back(host_ms[0]);
        t.unpack_host_ms.push_back(host_ms[1]);
      }
    }
    t.pass = verify();
    return t;
  };

  const char* names[3] = {"forall", "ordered", "unordered"};
  const double phase_bytes = 20.0 * (double)total * nvars;  // index + read + write per element
  std::printf("{\"grid\": [%d, %d, %d], \"halo\": %d, \"num_vars\": %d, \"num_cycles\": %d, "
              "\"var_size\": %d, \"list_total\": %lld, \"pack_hash\": %llu, \"unpack_hash\": %llu, "
              "\"loops_per_phase\": %d, \"phase_bytes\": %.0f",
              g[0], g[1], g[2], h, nvars, ncycles, var_size, (long long)total,
              (unsigned long long)list_hash(pl), (unsigned long long)list_hash(ul), kNeighbors * nvars,
              phase_bytes);
  bool all_ok = true;
  for (int w = 0; w < 3; ++w)
  {
    const Timing t = run_variant(w);
    all_ok = all_ok && t.pass;
    const double pk = median(t.pack_ms), up = median(t.unpack_ms);
    std::printf(", \"%s\": {\"pass\": %s, \"min_cycle_ms\": %.4f, \"median_cycle_ms\": %.4f, "
                "\"pack_device_ms\": %.4f, \"unpack_device_ms\": %.4f, \"pack_host_ms\": %.4f, "
                "\"unpack_host_ms\": %.4f, \"pack_GBs\": %.1f, \"unpack_GBs\": %.1f}",
                names[w], t.pass ? "true" : "false",
                1e3 * *std::min_element(t.cycle_s.begin(), t.cycle_s.end()), 1e3 * median(t.cycle_s),
                pk, up, median(t.pack_host_ms), median(t.unpack_host_ms),
                phase_bytes / (pk * 1e-3) / 1e9, phase_bytes / (up * 1e-3) / 1e9);
  }
  std::printf("}\n");
  for (auto& p : vars) check(hipFree(p), "free");
  check(hipFree(d_pl), "free");
  check(hipFree(d_ul), "free");
  check(hipFree(d_buf), "free");
  return all_ok ? 0 : 3;
}
