// Front-end test driver: RAJA API calls with seq_exec / omp_parallel_for_exec
// and, when built by hipcc, the cdna4 policies, checked against each other
// and against closed forms, following the reference's test cases
// (test/functional/forall/reduce-basic, test/functional/scan,
// test/unit/algorithm, test/functional/workgroup, test/unit/resource).
// Prints one line per check and exits with the number of failures.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "RAJA/RAJA.hpp"

static int g_fail = 0;
static int g_pass = 0;

#define CHECK(cond, name)                                                     \
  do                                                                          \
  {                                                                           \
    if (cond)                                                                 \
    {                                                                         \
      ++g_pass;                                                               \
    }                                                                         \
    else                                                                      \
    {                                                                         \
      ++g_fail;                                                               \
      std::printf("FAIL %s (%s:%d)\n", name, __FILE__, __LINE__);             \
    }                                                                         \
  } while (0)

// splitmix64 of (seed << 40) + i: the same generator the library uses
static uint64_t mix(uint64_t seed, uint64_t i)
{
  uint64_t z = (seed << 40) + i + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---- host policies ------------------------------------------------------------

template<typename EXEC, typename REDUCE>
void test_host_reducers(const char* tag)
{
  const int N = 2057;
  std::vector<int> a(N);
  std::vector<double> d(N);
  for (int i = 0; i < N; ++i)
  {
    a[i] = (int)(mix(1, i) % 100);
    d[i] = (double)a[i];
  }
  const int planted = N * 2 / 3;
  a[planted] = -7;
  d[planted] = -7.0;
  long long ref_sum = 0;
  for (int v : a) ref_sum += v;

  int* ap = a.data();
  double* dp = d.data();
  RAJA::ReduceSum<REDUCE, int> sum(0), sum2(2);
  RAJA::ReduceMin<REDUCE, int> mn(std::numeric_limits<int>::max());
  RAJA::ReduceMax<REDUCE, double> mx(-1.0);
  RAJA::ReduceMinLoc<REDUCE, double> mloc(std::numeric_limits<double>::max(), -1);
  RAJA::ReduceMaxLoc<REDUCE, int> xloc(std::numeric_limits<int>::min(), -1);
  RAJA::ReduceBitOr<REDUCE, int> bor(0);
  RAJA::ReduceBitAnd<REDUCE, int> band(~0);
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=](RAJA::Index_type i) {
    sum += ap[i];
    sum2 += ap[i];
    mn.min(ap[i]);
    mx.max(dp[i]);
    mloc.minloc(dp[i], i);
    xloc.maxloc(ap[i], i);
    bor |= (ap[i] & 0xF);
    band &= (ap[i] | 0x100);
  });
  std::string t(tag);
  CHECK(sum.get() == ref_sum, (t + " ReduceSum").c_str());
  CHECK(sum2.get() == ref_sum + 2, (t + " ReduceSum init offset").c_str());
  CHECK(mn.get() == -7, (t + " ReduceMin").c_str());
  CHECK(mx.get() == 99.0, (t + " ReduceMax").c_str());
  CHECK(mloc.get() == -7.0 && mloc.getLoc() == planted, (t + " ReduceMinLoc").c_str());
  int first99 = -1;
  for (int i = 0; i < N; ++i)
    if (a[i] == 99)
    {
      first99 = i;
      break;
    }
  CHECK(xloc.get() == 99 && xloc.getLoc() == first99, (t + " ReduceMaxLoc first index").c_str());
  CHECK(bor.get() == 0xF, (t + " ReduceBitOr").c_str());
  // accumulation across two foralls, then reset
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=](RAJA::Index_type i) { sum += ap[i]; });
  CHECK(sum.get() == 2 * ref_sum, (t + " ReduceSum accumulates").c_str());
  sum.reset(5);
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, 10), [=](RAJA::Index_type i) { sum += 1; });
  CHECK(sum.get() == 15, (t + " ReduceSum reset").c_str());
}

template<typename EXEC>
void test_host_params(const char* tag)
{
  const int N = 1000;
  std::vector<int> a(N);
  for (int i = 0; i < N; ++i) a[i] = (i % 2 == 0) ? 1 : -1;
  a[0] = 3;
  a[N / 2] = -100;
  a[N / 2 + 1] = 100;
  int* ap = a.data();
  int s = 0, mn = std::numeric_limits<int>::max(), mx = std::numeric_limits<int>::min();
  RAJA::expt::ValLoc<int> ml(std::numeric_limits<int>::max(), -1);
  int m2 = std::numeric_limits<int>::max();
  RAJA::Index_type l2 = -1;
  using VSUM = RAJA::expt::ValOp<int, RAJA::operators::plus>;
  using VMIN = RAJA::expt::ValOp<int, RAJA::operators::minimum>;
  using VMAX = RAJA::expt::ValOp<int, RAJA::operators::maximum>;
  using VLMIN = RAJA::expt::ValLocOp<int, RAJA::Index_type, RAJA::operators::minimum>;
  RAJA::forall<EXEC>(
      RAJA::RangeSegment(0, N), RAJA::expt::Reduce<RAJA::operators::plus>(&s),
      RAJA::expt::Reduce<RAJA::operators::minimum>(&mn),
      RAJA::expt::Reduce<RAJA::operators::maximum>(&mx),
      RAJA::expt::Reduce<RAJA::operators::minimum>(&ml), RAJA::Name("params"),
      RAJA::expt::ReduceLoc<RAJA::operators::minimum>(&m2, &l2),
      [=](RAJA::Index_type i, VSUM& vs, VMIN& vmn, VMAX& vmx, VLMIN& vml, VLMIN& vml2) {
        vs += ap[i];
        vmn.min(ap[i]);
        vmx.max(ap[i]);
        vml.minloc(ap[i], i);
        vml2.minloc(ap[i], i);
      });
  int ref = 0;
  for (int v : a) ref += v;
  std::string t(tag);
  CHECK(s == ref, (t + " expt sum").c_str());
  CHECK(mn == -100 && mx == 100, (t + " expt min/max").c_str());
  CHECK(ml.getVal() == -100 && ml.getLoc() == N / 2, (t + " expt ValLoc").c_str());
  CHECK(m2 == -100 && l2 == N / 2, (t + " expt ReduceLoc").c_str());
}

template<typename EXEC>
void test_host_scans_sorts(const char* tag)
{
  std::string t(tag);
  for (int n : {0, 357, 32000})
  {
    std::vector<long> in(n), out(n);
    std::iota(in.begin(), in.end(), 1);
    RAJA::inclusive_scan<EXEC>(RAJA::make_span(in.data(), n), RAJA::make_span(out.data(), n));
    bool ok = true;
    for (long i = 0; i < n; ++i) ok = ok && out[i] == (i + 1) * (i + 2) / 2;
    CHECK(ok, (t + " inclusive_scan iota").c_str());
    RAJA::exclusive_scan<EXEC>(RAJA::make_span(in.data(), n), RAJA::make_span(out.data(), n),
                               RAJA::operators::plus<long>{}, 13L);
    ok = true;
    for (long i = 0; i < n; ++i) ok = ok && out[i] == 13 + i * (i + 1) / 2;
    CHECK(ok, (t + " exclusive_scan iota offset 13").c_str());
    std::vector<long> ip(in);
    RAJA::exclusive_scan_inplace<EXEC>(RAJA::make_span(ip.data(), n), RAJA::operators::maximum<long>{});
    ok = true;
    for (long i = 1; i < n; ++i) ok = ok && ip[i] == i;
    CHECK(ok, (t + " exclusive_scan_inplace max").c_str());
  }
  const int n = 10000;
  std::vector<int> k(n), v(n);
  for (int i = 0; i < n; ++i)
  {
    k[i] = (int)(mix(3, i) % (2 * n + 1)) - n;
    v[i] = i;
  }
  std::vector<int> ks(k), vs(v);
  RAJA::stable_sort_pairs<EXEC>(RAJA::make_span(ks.data(), n), RAJA::make_span(vs.data(), n));
  bool ok = true;
  for (int i = 1; i < n; ++i)
    ok = ok && (ks[i - 1] < ks[i] || (ks[i - 1] == ks[i] && vs[i - 1] < vs[i]));
  for (int i = 0; i < n; ++i) ok = ok && k[vs[i]] == ks[i];
  CHECK(ok, (t + " stable_sort_pairs").c_str());
  std::vector<int> kd(k);
  RAJA::sort<EXEC>(RAJA::make_span(kd.data(), n), RAJA::operators::greater<int>{});
  ok = true;
  for (int i = 1; i < n; ++i) ok = ok && kd[i - 1] >= kd[i];
  CHECK(ok, (t + " sort greater").c_str());
}

template<typename WORK>
void test_host_workgroup(const char* tag)
{
  using Pol = RAJA::WorkGroupPolicy<WORK, RAJA::ordered, RAJA::constant_stride_array_of_objects>;
  using Pool = RAJA::WorkPool<Pol, RAJA::Index_type, RAJA::xargs<int>, std::allocator<char>>;
  const int N = 4096;
  std::vector<int> a(N, -1);
  int* ap = a.data();
  Pool pool;
  std::vector<std::pair<int, int>> ranges = {{0, 100}, {100, 100}, {150, 1024}, {2000, 4096}};
  for (auto r : ranges)
  {
    pool.enqueue(RAJA::RangeSegment(r.first, r.second), [=](RAJA::Index_type i, int off) { ap[i] = (int)i + off; });
  }
  auto group = pool.instantiate();
  CHECK(pool.num_loops() == 0 && group.num_loops() == 3, (std::string(tag) + " workgroup enqueue/instantiate").c_str());
  group.run(5);
  bool ok = true;
  for (int i = 0; i < N; ++i)
  {
    const bool inside = (i < 100) || (i >= 150 && i < 1024) || (i >= 2000);
    ok = ok && (inside ? a[i] == i + 5 : a[i] == -1);
  }
  CHECK(ok, (std::string(tag) + " workgroup i+5").c_str());
}

// ---- cdna4 --------------------------------------------------------------------------

// ---- multi (binned) reducers (test/functional/forall/multi-reduce-basic) ----------

template<typename EXEC, typename MPOL, typename A, typename D>
void run_multi_cases(const std::string& t, int N, int nb, A a, D x, const std::vector<int>& ha,
                     const std::vector<double>& hx)
{
  std::vector<long long> init(nb);
  for (int b = 0; b < nb; ++b) init[b] = b;
  RAJA::MultiReduceSum<MPOL, long long> isum(init);
  RAJA::MultiReduceSum<MPOL, double> dsum(nb);
  RAJA::MultiReduceMin<MPOL, double> dmin(nb);
  RAJA::MultiReduceMax<MPOL, int> imax(nb, -5000);
  RAJA::MultiReduceBitOr<MPOL, int> bor(nb);
  RAJA::MultiReduceBitAnd<MPOL, int> band(nb);
  for (int rep = 0; rep < 2; ++rep)  // results accumulate across loops
  {
    RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=] RAJA_HOST_DEVICE(RAJA::Index_type i) {
      const int b = (int)(i % nb);
      isum[b] += a[i];
      dsum[b] += x[i];
      dmin[b].min(x[i]);
      imax[b].max(a[i]);
      bor[b] |= (a[i] & 0xFF);
      band[b] &= (a[i] | 0x100);
    });
  }
  std::vector<long long> ris(nb);
  std::vector<double> rds(nb), rmin(nb, std::numeric_limits<double>::max());
  std::vector<int> rmax(nb, -5000), ror(nb, 0), rand_(nb, ~0);
  std::vector<double> abs_sum(nb, 0.0);
  for (int b = 0; b < nb; ++b) ris[b] = b;
  for (int rep = 0; rep < 2; ++rep)
  {
    for (int i = 0; i < N; ++i)
    {
      const int b = i % nb;
      ris[b] += ha[i];
      rds[b] += hx[i];
      abs_sum[b] += std::fabs(hx[i]);
      rmin[b] = std::min(rmin[b], hx[i]);
      rmax[b] = std::max(rmax[b], ha[i]);
      ror[b] |= (ha[i] & 0xFF);
      rand_[b] &= (ha[i] | 0x100);
    }
  }
  bool ok_is = true, ok_ds = true, ok_min = true, ok_max = true, ok_bits = true;
  for (int b = 0; b < nb; ++b)
  {
    ok_is = ok_is && isum.get(b) == ris[b];
    ok_ds = ok_ds && std::fabs(dsum.get(b) - rds[b]) <= 64 * 2.2e-16 * abs_sum[b] + 1e-300;
    ok_min = ok_min && dmin.get(b) == rmin[b];
    ok_max = ok_max && imax.get(b) == rmax[b];
    ok_bits = ok_bits && bor.get(b) == ror[b] && band.get(b) == rand_[b];
  }
  CHECK(ok_is, (t + " MultiReduceSum int64 with init container, two loops").c_str());
  CHECK(ok_ds, (t + " MultiReduceSum double within bound").c_str());
  CHECK(ok_min, (t + " MultiReduceMin double").c_str());
  CHECK(ok_max, (t + " MultiReduceMax int with init").c_str());
  CHECK(ok_bits, (t + " MultiReduceBitOr / BitAnd").c_str());
  std::vector<long long> all(nb);
  isum.get_all(all);
  CHECK(all == ris && isum.size() == (size_t)nb, (t + " MultiReduce get_all / size").c_str());
  isum.reset();
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=] RAJA_HOST_DEVICE(RAJA::Index_type i) { isum[0] += 1; });
  CHECK(isum.get(0) == N && isum.get(1) == 0, (t + " MultiReduce reset to identity").c_str());
  isum.reset(3, 7);
  CHECK(isum.size() == 3 && isum.get(2) == 7, (t + " MultiReduce reset(num_bins, init)").c_str());
}

template<typename EXEC, typename MPOL>
void test_host_multi(const char* tag)
{
  const int N = 10007, nb = 37;
  std::vector<int> a(N);
  std::vector<double> x(N);
  for (int i = 0; i < N; ++i)
  {
    a[i] = (int)(mix(21, i) % 2000) - 1000;
    x[i] = (double)((int)(mix(22, i) % 20001) - 10000) * 0.125;
  }
  run_multi_cases<EXEC, MPOL>(tag, N, nb, a.data(), x.data(), a, x);
}

// ---- index sets (test/functional/indexset-build, forall/indexset) ---------------

// Range [0, 1000), a list of every 7th index in [1000, 3000), a stride range
// [3000, 6000) step 5, a range [6000, 6001).
template<typename IS>
std::vector<RAJA::Index_type> fill_index_set(IS& iset, const std::vector<RAJA::Index_type>& list_host,
                                             const RAJA::TypedListSegment<RAJA::Index_type>& list)
{
  iset.push_back(RAJA::RangeSegment(0, 1000));
  iset.push_back(list);
  iset.push_back(RAJA::RangeStrideSegment(3000, 6000, 5));
  iset.push_back(RAJA::RangeSegment(6000, 6001));
  std::vector<RAJA::Index_type> all;
  for (RAJA::Index_type i = 0; i < 1000; ++i) all.push_back(i);
  all.insert(all.end(), list_host.begin(), list_host.end());
  for (RAJA::Index_type i = 3000; i < 6000; i += 5) all.push_back(i);
  all.push_back(6000);
  return all;
}

template<typename SEGIT, typename EXEC, typename REDUCE>
void test_host_index_set(const char* tag)
{
  std::vector<RAJA::Index_type> lh;
  for (RAJA::Index_type i = 1000; i < 3000; i += 7) lh.push_back(i);
  RAJA::TypedListSegment<RAJA::Index_type> list(lh, RAJA::resources::Host::get_default());
  RAJA::TypedIndexSet<RAJA::RangeSegment, RAJA::TypedListSegment<RAJA::Index_type>, RAJA::RangeStrideSegment> iset;
  const auto all = fill_index_set(iset, lh, list);
  std::vector<int> hits(6001, 0);
  int* hp = hits.data();
  RAJA::ReduceSum<REDUCE, long long> s(0);
  RAJA::forall<RAJA::ExecPolicy<SEGIT, EXEC>>(iset, [=](RAJA::Index_type i) {
    hp[i] += 1;
    s += i;
  });
  bool ok = iset.getNumSegments() == 4 && iset.getLength() == all.size();
  long long ref = 0;
  std::vector<int> want(6001, 0);
  for (auto i : all)
  {
    want[i] = 1;
    ref += i;
  }
  ok = ok && hits == want && s.get() == ref;
  CHECK(ok, (std::string(tag) + " forall over a TypedIndexSet (range, list, stride)").c_str());
}

// ---- tool plugins (test/integration/plugin in the reference) ------------------

struct NameRecorder : public RAJA::util::PluginStrategy
{
  static std::vector<std::string>& names()
  {
    static std::vector<std::string> v;
    return v;
  }
  static std::vector<RAJA::Platform>& platforms()
  {
    static std::vector<RAJA::Platform> v;
    return v;
  }
  static long& posts()
  {
    static long n = 0;
    return n;
  }
  void preLaunch(const RAJA::util::PluginContext& p) override
  {
    names().push_back(p.kernel_name);
    platforms().push_back(p.platform);
  }
  void postLaunch(const RAJA::util::PluginContext&) override { ++posts(); }
};
static RAJA::util::PluginRegistry::add<NameRecorder> g_recorder("name-recorder", "records kernel names");

template<typename EXEC>
void test_plugins(const char* tag, RAJA::Platform want)
{
  const size_t before = NameRecorder::names().size();
  const long posts_before = NameRecorder::posts();
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, 10), RAJA::Name("plugin-named-loop"), [=] RAJA_HOST_DEVICE(RAJA::Index_type) {});
  RAJA::forall<EXEC>(RAJA::RangeSegment(0, 10), [=] RAJA_HOST_DEVICE(RAJA::Index_type) {});
  const auto& n = NameRecorder::names();
  const bool ok = n.size() == before + 2 && n[before] == "plugin-named-loop" && n[before + 1].empty() &&
                  NameRecorder::platforms()[before] == want && NameRecorder::posts() == posts_before + 2;
  CHECK(ok, (std::string(tag) + " plugins see forall with name and platform").c_str());
}

// ---- RAJA::launch cases (test/functional/launch in the reference) ---------------
// One routine runs on the host (ExecPlace::HOST, seq policies, host
// pointers) and on the device (ExecPlace::DEVICE, cdna4 policies, device
// pointers); the results are compared with closed forms on the host.

struct LaunchResults
{
  long long sum = 0, sum2 = 2;
  RAJA::expt::ValLoc<double> mloc{std::numeric_limits<double>::max(), -1};
  long long msum = 0;
  long long empty_sum = 5;
};

template<typename LPOL, typename GLOBAL, typename BLOCKS, typename TX, typename TY>
LaunchResults run_launch_cases(RAJA::ExecPlace place, int N, const int* a, const double* x,
                               int R, int D, int C, const int* m, int* m2, double* team_out,
                               double* y, long long* z, int nteams1d)
{
  LaunchResults out;
  using VSUM = RAJA::expt::ValOp<long long, RAJA::operators::plus>;
  using VLMIN = RAJA::expt::ValLocOp<double, RAJA::Index_type, RAJA::operators::minimum>;
  // (a) global-thread loop with three parameter reducers
  RAJA::launch<LPOL>(place, RAJA::LaunchParams(RAJA::Teams(nteams1d), RAJA::Threads(256)),
                     RAJA::expt::Reduce<RAJA::operators::plus>(&out.sum),
                     RAJA::expt::Reduce<RAJA::operators::plus>(&out.sum2),
                     RAJA::expt::Reduce<RAJA::operators::minimum>(&out.mloc), RAJA::Name("launch-a"),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx, VSUM& s, VSUM& s2, VLMIN& ml) {
                       RAJA::loop<GLOBAL>(ctx, RAJA::RangeSegment(0, N), [&](RAJA::Index_type i) {
                         s += a[i];
                         s2 += a[i];
                         ml.minloc(x[i], i);
                       });
                     });
  // (b) three nested loops over a 70-thread (10 x 7) team, with a reducer
  RAJA::launch<LPOL>(place, RAJA::LaunchParams(RAJA::Teams(16), RAJA::Threads(10, 7)),
                     RAJA::expt::Reduce<RAJA::operators::plus>(&out.msum),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx, VSUM& s) {
                       RAJA::loop<BLOCKS>(ctx, RAJA::RangeSegment(0, R), [&](RAJA::Index_type r) {
                         RAJA::loop<TY>(ctx, RAJA::RangeSegment(0, D), [&](RAJA::Index_type d) {
                           RAJA::loop<TX>(ctx, RAJA::RangeSegment(0, C), [&](RAJA::Index_type c) {
                             const RAJA::Index_type k = (r * D + d) * C + c;
                             m2[k] = 2 * m[k];
                             s += m[k];
                           });
                         });
                       });
                     });
  // (c) team shared memory and teamSync: team t sums x[128 t .. 128 t + 127]
  const int T = (N + 127) / 128;
  RAJA::launch<LPOL>(place, RAJA::LaunchParams(RAJA::Teams(T), RAJA::Threads(128), 128 * sizeof(double)),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx) {
                       RAJA::loop<BLOCKS>(ctx, RAJA::RangeSegment(0, T), [&](RAJA::Index_type t) {
                         double* sh = ctx.getSharedMemory<double>(128);
                         RAJA::loop<TX>(ctx, RAJA::RangeSegment(0, 128), [&](RAJA::Index_type k) {
                           const RAJA::Index_type i = t * 128 + k;
                           sh[k] = (i < N) ? x[i] : 0.0;
                         });
                         ctx.teamSync();
                         RAJA::loop<TX>(ctx, RAJA::RangeSegment(0, 1), [&](RAJA::Index_type) {
                           double acc = 0.0;
                           for (int k = 0; k < 128; ++k) acc += sh[k];
                           team_out[t] = acc;
                         });
                         ctx.teamSync();
                         ctx.releaseSharedMemory();
                       });
                     });
  // (d) tile + loop_icount, tile_tcount
  RAJA::launch<LPOL>(place, RAJA::LaunchParams(RAJA::Teams(64), RAJA::Threads(256)),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx) {
                       RAJA::tile<BLOCKS>(ctx, 256, RAJA::RangeSegment(0, N), [&](RAJA::RangeSegment const& ts) {
                         RAJA::loop_icount<TX>(ctx, ts, [&](RAJA::Index_type i, RAJA::Index_type k) {
                           y[i] = x[i] + (double)k;
                         });
                       });
                       RAJA::tile_tcount<BLOCKS>(ctx, 256, RAJA::RangeSegment(0, N),
                                                 [&](RAJA::RangeSegment const& ts, RAJA::Index_type tix) {
                                                   RAJA::loop<TX>(ctx, ts, [&](RAJA::Index_type i) { z[i] = tix; });
                                                 });
                     });
  // (e) no teams: the parameter keeps its target value
  RAJA::launch<LPOL>(place, RAJA::LaunchParams(RAJA::Teams(0), RAJA::Threads(64)),
                     RAJA::expt::Reduce<RAJA::operators::plus>(&out.empty_sum),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx, VSUM& s) {
                       RAJA::loop<GLOBAL>(ctx, RAJA::RangeSegment(0, N), [&](RAJA::Index_type i) { s += a[i]; });
                     });
  return out;
}

void check_launch_results(const std::string& t, bool device, const LaunchResults& got, int N,
                          const std::vector<int>& a,
                          const std::vector<double>& x, const std::vector<int>& m,
                          const std::vector<int>& m2, const std::vector<double>& team_out,
                          const std::vector<double>& y, const std::vector<long long>& z)
{
  long long ref = 0;
  for (int v : a) ref += v;
  const int im = (int)(std::min_element(x.begin(), x.end()) - x.begin());
  long long mref = 0;
  bool m2ok = true;
  for (size_t k = 0; k < m.size(); ++k)
  {
    mref += m[k];
    m2ok = m2ok && m2[k] == 2 * m[k];
  }
  bool tok = true;
  for (size_t t2 = 0; t2 < team_out.size(); ++t2)
  {
    double acc = 0.0;
    for (int k = 0; k < 128; ++k)
    {
      const size_t i = t2 * 128 + k;
      acc += (i < (size_t)N) ? x[i] : 0.0;
    }
    tok = tok && acc == team_out[t2];
  }
  bool yok = true, zok = true;
  for (int i = 0; i < N; ++i)
  {
    yok = yok && y[i] == x[i] + (double)(i % 256);
    zok = zok && z[i] == i / 256;
  }
  CHECK(got.sum == ref, (t + " launch expt sum").c_str());
  CHECK(got.sum2 == ref + 2, (t + " launch expt sum init offset").c_str());
  CHECK(got.mloc.getVal() == x[im] && got.mloc.getLoc() == im, (t + " launch expt minloc").c_str());
  CHECK(got.msum == mref && m2ok, (t + " launch nested loops, 70-thread teams").c_str());
  CHECK(tok, (t + " launch shared memory + teamSync").c_str());
  CHECK(yok, (t + " launch tile + loop_icount").c_str());
  CHECK(zok, (t + " launch tile_tcount").c_str());
  // a device launch with no teams runs nothing; the sequential launch runs
  // its body once whatever the team count (include/RAJA/policy/sequential/launch.hpp)
  CHECK(got.empty_sum == (device ? 5 : 5 + ref), (t + " launch with no teams").c_str());
}

struct LaunchData
{
  int N = 100003, R = 50, D = 13, C = 77;
  std::vector<int> a, m;
  std::vector<double> x;
  LaunchData()
  {
    a.resize(N);
    x.resize(N);
    m.resize((size_t)R * D * C);
    for (int i = 0; i < N; ++i)
    {
      a[i] = (int)(mix(11, i) % 1000) - 500;
      x[i] = (double)((int)(mix(12, i) % 2001) - 1000) * 0.5;  // ties: first index wins
    }
    for (size_t k = 0; k < m.size(); ++k) m[k] = (int)(mix(13, k) % 100);
  }
};

void test_host_launch()
{
  LaunchData L;
  using LP = RAJA::LaunchPolicy<RAJA::seq_launch_t>;
  using S = RAJA::LoopPolicy<RAJA::seq_exec>;
  std::vector<int> m2(L.m.size(), 0);
  const int T = (L.N + 127) / 128;
  std::vector<double> team_out(T), y(L.N);
  std::vector<long long> z(L.N);
  LaunchResults r = run_launch_cases<LP, S, S, S, S>(RAJA::ExecPlace::HOST, L.N, L.a.data(), L.x.data(), L.R,
                                                     L.D, L.C, L.m.data(), m2.data(), team_out.data(),
                                                     y.data(), z.data(), (L.N + 255) / 256);
  check_launch_results("seq", false, r, L.N, L.a, L.x, L.m, m2, team_out, y, z);
}

#if defined(RAJA_ENABLE_CDNA4)

template<typename T>
T* to_device(const std::vector<T>& h)
{
  RAJA::resources::Cdna4 r = RAJA::resources::Cdna4::get_default();
  T* d = r.allocate<T>(h.size() + 1);
  r.memcpy(d, h.data(), sizeof(T) * h.size());
  r.wait();
  return d;
}

template<typename T>
std::vector<T> to_host(const T* d, size_t n)
{
  std::vector<T> h(n);
  RAJA::resources::Cdna4 r = RAJA::resources::Cdna4::get_default();
  r.memcpy(h.data(), d, sizeof(T) * n);
  r.wait();
  return h;
}

template<typename T>
void dfree(T* p)
{
  RAJA::resources::Cdna4::get_default().deallocate(p);
}

void test_cdna4_forall_reducers()
{
  using EXEC = RAJA::cdna4_exec<256>;
  using RED = RAJA::cdna4_reduce;
  for (int N : {1, 28, 642, 2057, 100003, 1 << 20})
  {
    std::vector<int> a(N);
    std::vector<double> d(N);
    for (int i = 0; i < N; ++i)
    {
      a[i] = (int)(mix(1, i) % 100);
      d[i] = (double)((int)(mix(2, i) % 5));  // heavy ties
    }
    const int planted = N * 2 / 3;
    a[planted] = -7;
    long long ref = 0;
    for (int v : a) ref += v;
    int first_min_d = (int)(std::min_element(d.begin(), d.end()) - d.begin());
    int* ap = to_device(a);
    double* dp = to_device(d);
    RAJA::ReduceSum<RED, long long> sum(0), sum2(2);
    RAJA::ReduceMin<RED, int> mn;
    RAJA::ReduceMax<RED, int> mx;
    RAJA::ReduceMinLoc<RED, double> ml;
    RAJA::ReduceMaxLoc<RED, int> xl;
    RAJA::ReduceBitOr<RED, int> bo(0);
    RAJA::ReduceBitAnd<RED, int> ba(~0);
    RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
      sum += ap[i];
      sum2 += ap[i];
      mn.min(ap[i]);
      mx.max(ap[i]);
      ml.minloc(dp[i], i);
      xl.maxloc(ap[i], i);
      bo |= (ap[i] & 0xF);
      ba &= (ap[i] | 0x100);
    });
    int first_max_a = (int)(std::max_element(a.begin(), a.end()) - a.begin());
    int ref_and = ~0, ref_or = 0;
    for (int v : a)
    {
      ref_and &= (v | 0x100);
      ref_or |= (v & 0xF);
    }
    CHECK(sum.get() == ref, "cdna4 ReduceSum");
    CHECK(sum2.get() == ref + 2, "cdna4 ReduceSum init offset");
    CHECK(mn.get() == *std::min_element(a.begin(), a.end()), "cdna4 ReduceMin");
    CHECK(mx.get() == *std::max_element(a.begin(), a.end()), "cdna4 ReduceMax");
    CHECK(ml.get() == d[first_min_d] && ml.getLoc() == first_min_d, "cdna4 ReduceMinLoc first index");
    CHECK(xl.get() == a[first_max_a] && xl.getLoc() == first_max_a, "cdna4 ReduceMaxLoc first index");
    CHECK(bo.get() == ref_or && ba.get() == ref_and, "cdna4 ReduceBitOr/BitAnd");
    RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) { sum += ap[i]; });
    CHECK(sum.get() == 2 * ref, "cdna4 ReduceSum accumulates across foralls");
    sum.reset(5);
    RAJA::forall<EXEC>(RAJA::RangeSegment(0, 10), [=] RAJA_DEVICE(RAJA::Index_type) { sum += 1; });
    CHECK(sum.get() == 15, "cdna4 ReduceSum reset");
    RAJA::ReduceMinLoc<RED, double> init_dom(-1.0, -1);
    RAJA::forall<EXEC>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) { init_dom.minloc(dp[i], i); });
    CHECK(init_dom.get() == -1.0 && init_dom.getLoc() == -1, "cdna4 MinLoc init dominates");
    dfree(ap);
    dfree(dp);
  }
}

void test_cdna4_daxpy_and_params()
{
  const int N = 1 << 20;
  std::vector<double> x(N), y(N);
  for (int i = 0; i < N; ++i)
  {
    x[i] = (double)(mix(0, i) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    y[i] = (double)(mix(7, i) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
  double* xp = to_device(x);
  double* yp = to_device(y);
  const double a = 2.5;
  RAJA::forall<RAJA::cdna4_exec<256>>(RAJA::RangeSegment(0, N),
                                      [=] RAJA_DEVICE(RAJA::Index_type i) { yp[i] = a * xp[i] + yp[i]; });
  std::vector<double> yr(y);
  for (int i = 0; i < N; ++i) yr[i] = a * x[i] + y[i];
  std::vector<double> yh = to_host(yp, N);
  CHECK(std::memcmp(yh.data(), yr.data(), sizeof(double) * N) == 0, "cdna4 daxpy bit-exact (no contraction)");
  double dot = 0.0, mn = std::numeric_limits<double>::max();
  RAJA::expt::ValLoc<double> ml(std::numeric_limits<double>::max(), -1);
  using VSUM = RAJA::expt::ValOp<double, RAJA::operators::plus>;
  using VMIN = RAJA::expt::ValOp<double, RAJA::operators::minimum>;
  using VLMIN = RAJA::expt::ValLocOp<double, RAJA::Index_type, RAJA::operators::minimum>;
  RAJA::forall<RAJA::cdna4_exec<256>>(
      RAJA::RangeSegment(0, N), RAJA::expt::Reduce<RAJA::operators::plus>(&dot),
      RAJA::expt::Reduce<RAJA::operators::minimum>(&mn),
      RAJA::expt::Reduce<RAJA::operators::minimum>(&ml),
      [=] RAJA_DEVICE(RAJA::Index_type i, VSUM& s, VMIN& m, VLMIN& l) {
        s += xp[i] * yp[i];
        m.min(xp[i]);
        l.minloc(xp[i], i);
      });
  double ref = 0.0, absum = 0.0;
  for (int i = 0; i < N; ++i)
  {
    ref += x[i] * yr[i];
    absum += std::fabs(x[i] * yr[i]);
  }
  const int im = (int)(std::min_element(x.begin(), x.end()) - x.begin());
  CHECK(std::fabs(dot - ref) <= 2 * 20 * 2.2e-16 * absum * 2, "cdna4 expt dot within bound");
  CHECK(mn == x[im], "cdna4 expt min");
  CHECK(ml.getVal() == x[im] && ml.getLoc() == im, "cdna4 expt ValLoc");
  // repeatability: same launch shape, same bits
  double dot2 = 0.0;
  RAJA::forall<RAJA::cdna4_exec<256>>(RAJA::RangeSegment(0, N), RAJA::expt::Reduce<RAJA::operators::plus>(&dot2),
                                      [=] RAJA_DEVICE(RAJA::Index_type i, VSUM& s) { s += xp[i] * yp[i]; });
  CHECK(dot2 == dot, "cdna4 expt sum deterministic");
  dfree(xp);
  dfree(yp);
}

void test_cdna4_segments_resources()
{
  const int N = 5000;
  std::vector<int> idx;
  for (int i = 0; i < N; i += 3) idx.push_back(N - 1 - i);
  RAJA::resources::Cdna4 res = RAJA::resources::Cdna4::get_default();
  RAJA::TypedListSegment<int> list(idx, res);
  std::vector<int> a(N, 0);
  int* ap = to_device(a);
  RAJA::ReduceSum<RAJA::cdna4_reduce, long> cnt(0);
  RAJA::forall<RAJA::cdna4_exec<128>>(list, [=] RAJA_DEVICE(int i) {
    ap[i] = 1;
    cnt += 1;
  });
  std::vector<int> h = to_host(ap, N);
  bool ok = true;
  for (int i = 0; i < N; ++i) ok = ok && h[i] == (((N - 1 - i) % 3 == 0) ? 1 : 0);
  CHECK(ok && cnt.get() == (long)idx.size(), "cdna4 ListSegment forall");
  RAJA::forall<RAJA::cdna4_exec<64>>(RAJA::RangeStrideSegment(3, 1029, 3), [=] RAJA_DEVICE(RAJA::Index_type i) { ap[i] = 7; });
  h = to_host(ap, N);
  ok = true;
  for (int i = 3; i < 1029; i += 3) ok = ok && h[i] == 7;
  CHECK(ok, "cdna4 RangeStrideSegment forall");
  // async resources with an event dependency (test-resource-Depends)
  RAJA::resources::Cdna4 r1, r2;
  std::vector<int> z(N, 0);
  int* d1 = to_device(z);
  RAJA::resources::Event e =
      RAJA::forall<RAJA::cdna4_exec_async<256>>(r1, RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) { d1[i] = (int)i * -1; });
  r2.wait_for(&e);
  RAJA::forall<RAJA::cdna4_exec_async<256>>(r2, RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) { d1[i] *= 2; });
  r2.wait();
  h = to_host(d1, N);
  ok = true;
  for (int i = 0; i < N; ++i) ok = ok && h[i] == -2 * i;
  CHECK(ok, "cdna4 async resources + event dependency");
  // empty range: no launch, reducer untouched
  RAJA::ReduceSum<RAJA::cdna4_reduce, int> e0(3);
  RAJA::forall<RAJA::cdna4_exec<256>>(RAJA::RangeSegment(5, 5), [=] RAJA_DEVICE(RAJA::Index_type) { e0 += 1; });
  CHECK(e0.get() == 3, "cdna4 empty range");
  RAJA::synchronize<RAJA::cdna4_synchronize>();
  dfree(ap);
  dfree(d1);
}

void test_cdna4_scans_sorts()
{
  using EXEC = RAJA::cdna4_exec<256>;
  for (int n : {0, 357, 32000, 1000003})
  {
    std::vector<int> in(n), out(n);
    for (int i = 0; i < n; ++i) in[i] = (int)(mix(3, i) % 16);
    int* din = to_device(in);
    int* dout = to_device(out);
    RAJA::exclusive_scan<EXEC>(RAJA::make_span(din, n), RAJA::make_span(dout, n));
    std::vector<int> ref(n);
    RAJA::exclusive_scan<RAJA::seq_exec>(RAJA::make_span(in.data(), n), RAJA::make_span(ref.data(), n));
    CHECK(to_host(dout, n) == ref, "cdna4 exclusive_scan == seq");
    RAJA::inclusive_scan<EXEC>(RAJA::make_span(din, n), RAJA::make_span(dout, n), RAJA::operators::maximum<int>{});
    RAJA::inclusive_scan<RAJA::seq_exec>(RAJA::make_span(in.data(), n), RAJA::make_span(ref.data(), n), RAJA::operators::maximum<int>{});
    CHECK(to_host(dout, n) == ref, "cdna4 inclusive_scan max == seq");
    RAJA::exclusive_scan_inplace<EXEC>(RAJA::make_span(din, n), RAJA::operators::plus<int>{}, 13);
    RAJA::exclusive_scan_inplace<RAJA::seq_exec>(RAJA::make_span(in.data(), n), RAJA::operators::plus<int>{}, 13);
    CHECK(to_host(din, n) == in, "cdna4 exclusive_scan_inplace offset == seq");
    dfree(din);
    dfree(dout);
  }
  for (int n : {0, 1, 10, 100, 1000, 10000, 300007})
  {
    std::vector<int> k(n), v(n);
    for (int i = 0; i < n; ++i)
    {
      k[i] = (int)(mix(4, i) % (2 * n + 1)) - n;
      v[i] = i;
    }
    int* dk = to_device(k);
    int* dv = to_device(v);
    RAJA::sort_pairs<EXEC>(RAJA::make_span(dk, n), RAJA::make_span(dv, n), RAJA::operators::greater<int>{});
    std::vector<int> rk(k), rv(v);
    RAJA::stable_sort_pairs<RAJA::seq_exec>(RAJA::make_span(rk.data(), n), RAJA::make_span(rv.data(), n),
                                            RAJA::operators::greater<int>{});
    CHECK(to_host(dk, n) == rk && to_host(dv, n) == rv, "cdna4 sort_pairs greater == seq stable");
    std::vector<double> kd(n);
    for (int i = 0; i < n; ++i) kd[i] = 0.5 * k[i];
    double* dkd = to_device(kd);
    RAJA::stable_sort<EXEC>(RAJA::make_span(dkd, n));
    std::vector<double> rd(kd);
    RAJA::stable_sort<RAJA::seq_exec>(RAJA::make_span(rd.data(), n));
    CHECK(to_host(dkd, n) == rd, "cdna4 stable_sort double == seq");
    dfree(dk);
    dfree(dv);
    dfree(dkd);
  }
}

template<typename ORDER, typename DISPATCH>
void test_cdna4_workgroup_case(const char* name)
{
  const int N = 50000;
  std::vector<int> a(N, -1);
  int* ap = to_device(a);
  auto body = [=] RAJA_DEVICE(RAJA::Index_type i, int off) { ap[i] = (int)i + off; };
  using Body = decltype(body);
  using Pol = RAJA::WorkGroupPolicy<RAJA::cdna4_work<256>, ORDER, RAJA::constant_stride_array_of_objects, DISPATCH>;
  using Pool = RAJA::WorkPool<Pol, RAJA::Index_type, RAJA::xargs<int>, std::allocator<char>>;
  (void)sizeof(Body);
  Pool pool;
  std::vector<int> inside(N, 0);
  for (int l = 0; l < 40; ++l)
  {
    int b = (int)(mix(9, 2 * l) % N), len = (int)(mix(9, 2 * l + 1) % 1025);
    int e = std::min(N, b + len);
    // keep ranges disjoint: skip overlaps
    bool clash = false;
    for (int i = b; i < e; ++i) clash = clash || inside[i];
    if (clash) continue;
    for (int i = b; i < e; ++i) inside[i] = 1;
    pool.enqueue(RAJA::RangeSegment(b, e), body);
  }
  auto group = pool.instantiate();
  auto site = group.run(5);
  RAJA::resources::Cdna4 r = site.get_resource();
  r.wait();
  std::vector<int> h = to_host(ap, N);
  bool ok = true;
  for (int i = 0; i < N; ++i) ok = ok && (inside[i] ? h[i] == i + 5 : h[i] == -1);
  CHECK(ok, name);
  // run again (tables reused)
  group.run(7);
  h = to_host(ap, N);
  ok = true;
  for (int i = 0; i < N; ++i) ok = ok && (inside[i] ? h[i] == i + 7 : h[i] == -1);
  CHECK(ok, name);
  dfree(ap);
}

struct PackBody
{
  double* buf;
  const double* var;
  const int* list;
  RAJA_HOST_DEVICE void operator()(RAJA::Index_type i) const { buf[i] = var[list[i]]; }
};
struct UnpackBody
{
  const double* buf;
  double* var;
  const int* list;
  RAJA_HOST_DEVICE void operator()(RAJA::Index_type i) const { var[list[i]] = buf[i]; }
};

void test_cdna4_workgroups()
{
  using IFC = RAJA::indirect_function_call_dispatch;
  test_cdna4_workgroup_case<RAJA::unordered_cdna4_loop_tiles, IFC>("cdna4 workgroup unordered indirect");
  test_cdna4_workgroup_case<RAJA::ordered, IFC>("cdna4 workgroup ordered indirect");
  test_cdna4_workgroup_case<RAJA::reverse_ordered, IFC>("cdna4 workgroup reverse indirect");
  // halo pack/unpack on a small grid with direct dispatch of two body types
  const int g = 16, h = 2, s = g + 2 * h, size = s * s * s;
  std::vector<double> var(size);
  for (int i = 0; i < size; ++i) var[i] = i;
  std::vector<int> pl, ul;
  std::vector<int> lens;
  const int dirs[26][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}, {-1, -1, 0},
                           {-1, 1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},
                           {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1},
                           {-1, 1, 1}, {1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {1, 1, 1}};
  for (int l = 0; l < 26; ++l)
  {
    int pr[3][2], ur[3][2];
    for (int a = 0; a < 3; ++a)
    {
      const int d = dirs[l][a];
      pr[a][0] = d == 0 ? h : (d < 0 ? h : g);
      pr[a][1] = d == 0 ? g + h : (d < 0 ? 2 * h : g + h);
      ur[a][0] = d == 0 ? h : (d < 0 ? 0 : g + h);
      ur[a][1] = d == 0 ? g + h : (d < 0 ? h : g + 2 * h);
    }
    int n = 0;
    for (int k = pr[2][0]; k < pr[2][1]; ++k)
      for (int j = pr[1][0]; j < pr[1][1]; ++j)
        for (int i = pr[0][0]; i < pr[0][1]; ++i, ++n) pl.push_back(i + j * s + k * s * s);
    for (int k = ur[2][0]; k < ur[2][1]; ++k)
      for (int j = ur[1][0]; j < ur[1][1]; ++j)
        for (int i = ur[0][0]; i < ur[0][1]; ++i) ul.push_back(i + j * s + k * s * s);
    lens.push_back(n);
  }
  double* dvar = to_device(var);
  int* dpl = to_device(pl);
  int* dul = to_device(ul);
  std::vector<double> zero(pl.size(), 0.0);
  double* dbuf = to_device(zero);
  using Disp = RAJA::direct_dispatch<camp::list<RAJA::RangeSegment, PackBody>, camp::list<RAJA::RangeSegment, UnpackBody>>;
  using Pol = RAJA::WorkGroupPolicy<RAJA::cdna4_work<256>, RAJA::unordered_cdna4_loop_tiles,
                                    RAJA::constant_stride_array_of_objects, Disp>;
  using Pool = RAJA::WorkPool<Pol, RAJA::Index_type, RAJA::xargs<>, std::allocator<char>>;
  Pool pack_pool, unpack_pool;
  int off = 0;
  for (int l = 0; l < 26; ++l)
  {
    pack_pool.enqueue(RAJA::RangeSegment(0, lens[l]), PackBody{dbuf + off, dvar, dpl + off});
    unpack_pool.enqueue(RAJA::RangeSegment(0, lens[l]), UnpackBody{dbuf + off, dvar, dul + off});
    off += lens[l];
  }
  auto pg = pack_pool.instantiate();
  auto ug = unpack_pool.instantiate();
  pg.run();
  ug.run();
  std::vector<double> res = to_host(dvar, size);
  std::vector<double> ref(var);
  for (size_t q = 0; q < pl.size(); ++q) ref[ul[q]] = var[pl[q]];
  CHECK(res == ref, "cdna4 workgroup halo pack/unpack (direct dispatch)");
  dfree(dvar);
  dfree(dpl);
  dfree(dul);
  dfree(dbuf);
}

// An unordered WorkGroup run captured into a HIP graph and replayed: the
// group's device tables are built by an eager run on the capture stream,
// so the capture holds the one tile kernel.  Each replay, after the inputs
// change eagerly, must give what an eager run gives.
struct ScaleBody
{
  const double* in;
  double* out;
  double a;
  RAJA_HOST_DEVICE void operator()(RAJA::Index_type i) const { out[i] = a * in[i] + 1.0; }
};

void test_cdna4_workgroup_graph()
{
  const int nloops = 40;
  std::vector<int> lens;
  int total = 0;
  for (int l = 0; l < nloops; ++l)
  {
    lens.push_back(l % 7 == 0 ? 70000 + 13 * l : 1 + 97 * l);
    total += lens.back();
  }
  std::vector<double> hin(total);
  for (int i = 0; i < total; ++i) hin[i] = 0.5 * i - 3.0;
  double* din = to_device(hin);
  std::vector<double> zero(total, 0.0);
  double* dout = to_device(zero);
  hipStream_t st = nullptr;
  CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess, "graph: stream");
  auto res = RAJA::resources::Cdna4::CdnaFromStream(st);
  using Disp = RAJA::direct_dispatch<camp::list<RAJA::RangeSegment, ScaleBody>>;
  using Pol = RAJA::WorkGroupPolicy<RAJA::cdna4_work_async<256>, RAJA::unordered_cdna4_loop_tiles,
                                    RAJA::constant_stride_array_of_objects, Disp>;
  using Pool = RAJA::WorkPool<Pol, RAJA::Index_type, RAJA::xargs<>, std::allocator<char>>;
  Pool pool;
  int off = 0;
  for (int l = 0; l < nloops; ++l)
  {
    pool.enqueue(RAJA::RangeSegment(0, lens[l]), ScaleBody{din + off, dout + off, 1.0 + l});
    off += lens[l];
  }
  auto group = pool.instantiate();
  group.run(res);  // eager: builds and uploads the device tables
  res.wait();
  auto expect = [&](const std::vector<double>& in) {
    std::vector<double> r(total);
    int o = 0;
    for (int l = 0; l < nloops; ++l)
    {
      for (int i = 0; i < lens[l]; ++i) r[o + i] = (1.0 + l) * in[o + i] + 1.0;
      o += lens[l];
    }
    return r;
  };
  CHECK(to_host(dout, total) == expect(hin), "cdna4 workgroup eager run before capture");
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
  group.run(res);
  ok = ok && hipStreamEndCapture(st, &graph) == hipSuccess;
  ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  CHECK(ok, "cdna4 workgroup run captured into a graph");
  bool all = ok;
  for (int rep = 0; ok && rep < 4; ++rep)
  {
    for (int i = 0; i < total; ++i) hin[i] = (double)((i * (rep + 3)) % 1001) - 500.0;
    CHECK(hipMemcpy(din, hin.data(), sizeof(double) * total, hipMemcpyHostToDevice) == hipSuccess, "graph: input");
    CHECK(hipMemset(dout, 0, sizeof(double) * total) == hipSuccess, "graph: clear");
    all = all && hipGraphLaunch(exec, st) == hipSuccess;
    all = all && hipStreamSynchronize(st) == hipSuccess;
    all = all && to_host(dout, total) == expect(hin);
  }
  CHECK(all, "cdna4 workgroup graph replays (4, inputs changed eagerly between them)");
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  group.clear();
  dfree(din);
  dfree(dout);
  (void)raja_cdna4_stream_release(st);
  (void)hipStreamDestroy(st);
}

void test_cdna4_launch()
{
  LaunchData L;
  using LP = RAJA::LaunchPolicy<RAJA::seq_launch_t, RAJA::cdna4_launch_t<false>>;
  using G = RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_global_thread_x>;
  using B = RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_block_x_loop>;
  using TX = RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_thread_x_loop>;
  using TY = RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_thread_y_loop>;
  const int T = (L.N + 127) / 128;
  int* da = to_device(L.a);
  double* dx = to_device(L.x);
  int* dm = to_device(L.m);
  int* dm2 = to_device(std::vector<int>(L.m.size(), 0));
  double* dteam = to_device(std::vector<double>(T, 0.0));
  double* dy = to_device(std::vector<double>(L.N, 0.0));
  long long* dz = to_device(std::vector<long long>(L.N, 0));
  LaunchResults r = run_launch_cases<LP, G, B, TX, TY>(RAJA::ExecPlace::DEVICE, L.N, da, dx, L.R, L.D, L.C, dm,
                                                       dm2, dteam, dy, dz, (L.N + 255) / 256);
  check_launch_results("cdna4", true, r, L.N, L.a, L.x, L.m, to_host(dm2, L.m.size()), to_host(dteam, T),
                       to_host(dy, L.N), to_host(dz, L.N));
  // the same policy list on the host place runs the seq entries on host data
  {
    std::vector<int> m2(L.m.size(), 0);
    std::vector<double> team_out(T), y(L.N);
    std::vector<long long> z(L.N);
    LaunchResults h = run_launch_cases<LP, G, B, TX, TY>(RAJA::ExecPlace::HOST, L.N, L.a.data(), L.x.data(),
                                                         L.R, L.D, L.C, L.m.data(), m2.data(),
                                                         team_out.data(), y.data(), z.data(),
                                                         (L.N + 255) / 256);
    check_launch_results("cdna4-list host place", false, h, L.N, L.a, L.x, L.m, m2, team_out, y, z);
  }
  // determinism: the same launch shape gives the same bits
  const int N = L.N;
  double s1 = 0.0, s2 = 0.0;
  using VD = RAJA::expt::ValOp<double, RAJA::operators::plus>;
  for (double* tgt : {&s1, &s2})
  {
    RAJA::launch<LP>(RAJA::ExecPlace::DEVICE, RAJA::LaunchParams(RAJA::Teams(97), RAJA::Threads(96)),
                     RAJA::expt::Reduce<RAJA::operators::plus>(tgt),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx, VD& s) {
                       RAJA::loop<RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_global_x_loop>>(
                           ctx, RAJA::RangeSegment(0, N), [&](RAJA::Index_type i) { s += dx[i] * 1e-3; });
                     });
  }
  CHECK(s1 == s2, "cdna4 launch FP sum deterministic (96-thread teams)");
  // old-style reducers captured by a launch body
  {
    RAJA::ReduceSum<RAJA::cdna4_reduce, long long> sum(0);
    RAJA::ReduceMinLoc<RAJA::cdna4_reduce, double> ml;
    RAJA::launch<LP>(RAJA::ExecPlace::DEVICE, RAJA::LaunchParams(RAJA::Teams(200), RAJA::Threads(100)),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx) {
                       RAJA::loop<RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_global_x_loop>>(
                           ctx, RAJA::RangeSegment(0, N), [&](RAJA::Index_type i) {
                             sum += da[i];
                             ml.minloc(dx[i], i);
                           });
                     });
    long long ref = 0;
    for (int v : L.a) ref += v;
    const int im = (int)(std::min_element(L.x.begin(), L.x.end()) - L.x.begin());
    CHECK(sum.get() == ref, "cdna4 launch old-style ReduceSum");
    CHECK(ml.get() == L.x[im] && ml.getLoc() == im, "cdna4 launch old-style ReduceMinLoc");
  }
  dfree(da);
  dfree(dx);
  dfree(dm);
  dfree(dm2);
  dfree(dteam);
  dfree(dy);
  dfree(dz);
}

void test_cdna4_multi()
{
  const int nb = 97;
  for (int N : {1000, 1 << 20})
  {
    std::vector<int> a(N);
    std::vector<double> x(N);
    for (int i = 0; i < N; ++i)
    {
      a[i] = (int)(mix(21, i) % 2000) - 1000;
      x[i] = (double)((int)(mix(22, i) % 20001) - 10000) * 0.125;
    }
    int* da = to_device(a);
    double* dx = to_device(x);
    run_multi_cases<RAJA::cdna4_exec<256>, RAJA::cdna4_multi_reduce>("cdna4 N=" + std::to_string(N), N, nb,
                                                                      (const int*)da, (const double*)dx, a, x);
    // inside RAJA::launch, 200 teams of 96 threads
    RAJA::MultiReduceSum<RAJA::cdna4_multi_reduce, long long> ls(nb);
    using LP = RAJA::LaunchPolicy<RAJA::seq_launch_t, RAJA::cdna4_launch_t<false>>;
    RAJA::launch<LP>(RAJA::ExecPlace::DEVICE, RAJA::LaunchParams(RAJA::Teams(200), RAJA::Threads(96)),
                     [=] RAJA_HOST_DEVICE(RAJA::LaunchContext ctx) {
                       RAJA::loop<RAJA::LoopPolicy<RAJA::seq_exec, RAJA::cdna4_global_x_loop>>(
                           ctx, RAJA::RangeSegment(0, N), [&](RAJA::Index_type i) { ls[i % nb] += da[i]; });
                     });
    bool ok = true;
    std::vector<long long> ref(nb, 0);
    for (int i = 0; i < N; ++i) ref[i % nb] += a[i];
    for (int b = 0; b < nb; ++b) ok = ok && ls.get(b) == ref[b];
    CHECK(ok, "cdna4 MultiReduceSum inside launch");
    dfree(da);
    dfree(dx);
  }
  // many bins: rows shared between blocks (agent-scope atomics), still exact for integers
  {
    const int N = 1 << 22, nbig = 1 << 20;
    std::vector<int> a(N);
    for (int i = 0; i < N; ++i) a[i] = (int)(mix(23, i) % 100);
    int* da = to_device(a);
    RAJA::MultiReduceSum<RAJA::cdna4_multi_reduce, long long> big(nbig);
    RAJA::forall<RAJA::cdna4_exec<256>>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
      big[(i * 7919) % nbig] += da[i];
    });
    std::vector<long long> ref(nbig, 0), got(nbig);
    for (int i = 0; i < N; ++i) ref[((long long)i * 7919) % nbig] += a[i];
    big.get_all(got);
    CHECK(got == ref, "cdna4 MultiReduceSum 2^20 bins (shared rows)");
    dfree(da);
  }
}

// Floating-point multi-reductions are deterministic: each wave folds its
// calls into a private row (same-bin lanes in lane order, calls in program
// order) and get() folds the rows in wave order.  The host below replays
// that order for the launch shape of a loop that reduces (tiles of
// kReduceItems*B indices, a grid of min(tiles, CUs * kReduceThreadsPerCU / B)
// blocks: 8*B and 512 threads per CU), so the device bits must equal it, for
// two block sizes, and repeats must give the same bits.
template<int B>
std::vector<double> emulate_wave_rows(const std::vector<double>& x, int nb, size_t (*bin)(long long))
{
  const long long N = (long long)x.size(), U = RAJA::cdna4::detail::kReduceItems;
  const long long per_block = (long long)B * U;
  const long long cap =
      (long long)raja_cdna4_num_cus() * std::max(1, RAJA::cdna4::detail::kReduceThreadsPerCU / B);
  const long long grid = std::min((N + per_block - 1) / per_block, cap);
  const int wpb = B / 64;
  std::vector<double> rows((size_t)(grid * wpb) * nb, 0.0);
  for (long long b = 0; b < grid; ++b)
  {
    for (int w = 0; w < wpb; ++w)
    {
      double* row = &rows[(size_t)(b * wpb + w) * nb];
      for (long long base = b * per_block; base < N; base += grid * per_block)
      {
        for (long long u = 0; u < U; ++u)
        {
          bool done[64] = {false};
          for (int l0 = 0; l0 < 64; ++l0)
          {
            const long long i0 = base + u * B + w * 64 + l0;
            if (i0 >= N || done[l0]) continue;
            const size_t lb = bin(i0);
            double r = 0.0;
            for (int l = l0; l < 64; ++l)
            {
              const long long i = base + u * B + w * 64 + l;
              if (i < N && !done[l] && bin(i) == lb)
              {
                r = r + x[(size_t)i];
                done[l] = true;
              }
            }
            row[lb] = row[lb] + r;
          }
        }
      }
    }
  }
  std::vector<double> out((size_t)nb);
  for (int k = 0; k < nb; ++k)
  {
    double r = 0.0;
    for (long long q = 0; q < grid * wpb; ++q) r = r + rows[(size_t)q * nb + k];
    out[(size_t)k] = 0.0 + (0.0 + r);  // host total (identity) then init (identity)
  }
  return out;
}

static size_t bin_spread(long long i) { return (size_t)((i * 919) % 1000); }
static size_t bin_runs(long long i) { return (size_t)((i / 3) % 7); }

template<int B>
void check_multi_deterministic(const std::vector<double>& x, const double* dx)
{
  const long long N = (long long)x.size();
  for (int which = 0; which < 2; ++which)
  {
    const int nb = which == 0 ? 1000 : 7;
    std::vector<double> got[2];
    for (int rep = 0; rep < 2; ++rep)
    {
      RAJA::MultiReduceSum<RAJA::cdna4_multi_reduce, double> s(nb);
      if (which == 0)
      {
        RAJA::forall<RAJA::cdna4_exec<B>>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
          s[(i * 919) % 1000] += dx[i];
        });
      }
      else
      {
        RAJA::forall<RAJA::cdna4_exec<B>>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
          s[(i / 3) % 7] += dx[i];
        });
      }
      got[rep].resize((size_t)nb);
      s.get_all(got[rep]);
    }
    const auto want = emulate_wave_rows<B>(x, nb, which == 0 ? bin_spread : bin_runs);
    const bool same = std::memcmp(got[0].data(), got[1].data(), sizeof(double) * nb) == 0;
    const bool exact = std::memcmp(got[0].data(), want.data(), sizeof(double) * nb) == 0;
    const std::string t = "cdna4 FP MultiReduceSum B=" + std::to_string(B) + (which == 0 ? " 1000 bins" : " 7 bins (runs)");
    CHECK(same, (t + ": repeats bit-identical").c_str());
    CHECK(exact, (t + ": bits of the documented fold order").c_str());
  }
}

void test_cdna4_multi_deterministic()
{
  const int N = (1 << 21) + 777;
  std::vector<double> x(N);
  for (int i = 0; i < N; ++i) x[i] = (double)(mix(31, i) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  double* dx = to_device(x);
  check_multi_deterministic<256>(x, dx);
  check_multi_deterministic<512>(x, dx);
  // min over FP bins: exact
  RAJA::MultiReduceMin<RAJA::cdna4_multi_reduce, double> mn(13);
  RAJA::forall<RAJA::cdna4_exec<256>>(RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
    mn[i % 13].min(dx[i]);
  });
  std::vector<double> ref(13, std::numeric_limits<double>::max());
  for (int i = 0; i < N; ++i) ref[i % 13] = std::min(ref[i % 13], x[i]);
  bool ok = true;
  for (int b = 0; b < 13; ++b) ok = ok && mn.get(b) == ref[b];
  CHECK(ok, "cdna4 FP MultiReduceMin exact");
  dfree(dx);
}

void test_cdna4_index_set()
{
  std::vector<RAJA::Index_type> lh;
  for (RAJA::Index_type i = 1000; i < 3000; i += 7) lh.push_back(i);
  RAJA::TypedListSegment<RAJA::Index_type> list(lh, RAJA::resources::Cdna4::get_default());
  RAJA::TypedIndexSet<RAJA::RangeSegment, RAJA::TypedListSegment<RAJA::Index_type>, RAJA::RangeStrideSegment> iset;
  const auto all = fill_index_set(iset, lh, list);
  int* hp = to_device(std::vector<int>(6001, 0));
  RAJA::ReduceSum<RAJA::cdna4_reduce, long long> s(0);
  RAJA::forall<RAJA::ExecPolicy<RAJA::seq_segit, RAJA::cdna4_exec<256>>>(iset, [=] RAJA_DEVICE(RAJA::Index_type i) {
    hp[i] += 1;
    s += i;
  });
  std::vector<int> want(6001, 0);
  long long ref = 0;
  for (auto i : all)
  {
    want[i] = 1;
    ref += i;
  }
  CHECK(to_host(hp, 6001) == want && s.get() == ref, "cdna4 forall over a TypedIndexSet (range, list, stride)");
  // segments run in order: segment 2 sees what segment 1 wrote
  RAJA::TypedIndexSet<RAJA::RangeSegment> two;
  two.push_back(RAJA::RangeSegment(0, 1));
  two.push_back(RAJA::RangeSegment(1, 4097));
  int* dp = to_device(std::vector<int>(4097, 0));
  RAJA::forall<RAJA::ExecPolicy<RAJA::seq_segit, RAJA::cdna4_exec<256>>>(two, [=] RAJA_DEVICE(RAJA::Index_type i) {
    dp[i] = (i == 0) ? 41 : dp[0] + 1;
  });
  const auto dh = to_host(dp, 4097);
  bool ok = dh[0] == 41;
  for (int i = 1; i < 4097; ++i) ok = ok && dh[i] == 42;
  CHECK(ok, "cdna4 index set segments run in order");
  dfree(hp);
  dfree(dp);
}

#endif

int main()
{
  test_host_reducers<RAJA::seq_exec, RAJA::seq_reduce>("seq");
  test_host_reducers<RAJA::omp_parallel_for_exec, RAJA::omp_reduce>("omp");
  test_host_params<RAJA::seq_exec>("seq");
  test_host_params<RAJA::omp_parallel_for_exec>("omp");
  test_host_scans_sorts<RAJA::seq_exec>("seq");
  test_host_scans_sorts<RAJA::omp_parallel_for_exec>("omp");
  test_host_workgroup<RAJA::seq_work>("seq");
  test_host_workgroup<RAJA::omp_work>("omp");
  test_host_launch();
  test_plugins<RAJA::seq_exec>("seq", RAJA::Platform::host);
  test_host_multi<RAJA::seq_exec, RAJA::seq_multi_reduce>("seq");
  test_host_multi<RAJA::omp_parallel_for_exec, RAJA::omp_multi_reduce>("omp");
  test_host_index_set<RAJA::seq_segit, RAJA::seq_exec, RAJA::seq_reduce>("seq_segit/seq");
  test_host_index_set<RAJA::omp_parallel_segit, RAJA::seq_exec, RAJA::omp_reduce>("omp_segit/seq");
#if defined(RAJA_ENABLE_CDNA4)
  test_cdna4_forall_reducers();
  test_cdna4_daxpy_and_params();
  test_cdna4_segments_resources();
  test_cdna4_scans_sorts();
  test_cdna4_workgroups();
  test_cdna4_workgroup_graph();
  test_cdna4_launch();
  test_plugins<RAJA::cdna4_exec<256>>("cdna4", RAJA::Platform::cdna4);
  test_cdna4_multi();
  test_cdna4_multi_deterministic();
  test_cdna4_index_set();
#endif
  RAJA::util::finalize_plugins();
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail;
}
