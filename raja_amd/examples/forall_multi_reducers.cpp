// One forall with four reducers -- ReduceSum, ReduceMin, ReduceMax and
// ReduceMinLoc -- over N FP64 values ~ N(0,1), through RAJA::forall on the
// cdna4 policies: the multi-reduction workload of the reference's reducer
// tests and docs (test/functional/forall/reduce-multiple-segment,
// docs/sphinx/user_guide/feature/reduction.rst:280-323,
// include/RAJA/policy/hip/reduce.hpp:960-1161).  One read of x, 8 B per
// element.
//
// The loop runs three times per repetition: with old-style reducers (get()
// synchronises), with four expt::Reduce parameters (resolved at the end of
// forall), and with old-style reducers and a body that reads x through a
// nontemporal load (a read-once hint outside RAJA's API, reported apart).
// x comes from the library's counter generator (normal, seed 5).  The host
// checks the last repetition: min, max and the first index of the minimum
// exactly, the sum within 2*ceil(log2 N)*eps*sum|x| of a compensated
// (Kahan-Neumaier) sum, and the three runs bit for bit.  Prints one JSON
// line.
//
// Usage: forall_multi_reducers [log2n] [reps]     (default 28 10)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <vector>

#include "RAJA/RAJA.hpp"
#include "raja_cdna4.h"

namespace
{

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
  {
    std::fprintf(stderr, "forall_multi_reducers: %s: %s\n", what, hipGetErrorString(e));
    std::exit(2);
  }
}

void check_lib(int st, const char* what)
{
  if (st != RAJA_CDNA4_SUCCESS)
  {
    std::fprintf(stderr, "forall_multi_reducers: %s: %s\n", what, raja_cdna4_last_error());
    std::exit(2);
  }
}

double median(std::vector<double> v)
{
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

struct Result
{
  double sum = 0, mn = 0, mx = 0, lv = 0;
  RAJA::Index_type loc = -1;
};

}  // namespace

int main(int argc, char** argv)
{
  const int log2n = argc > 1 ? std::atoi(argv[1]) : 28;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 10;
  if (log2n < 4 || log2n > 30 || reps < 1)
  {
    std::fprintf(stderr, "usage: %s [log2n 4..30] [reps >= 1]\n", argv[0]);
    return 1;
  }
  const RAJA::Index_type N = (RAJA::Index_type)1 << log2n;
  check_lib(raja_cdna4_init(0), "init");
  auto res = RAJA::resources::Cdna4::get_default();
  hipStream_t st = res.get_stream();
  double* x = nullptr;
  check(hipMalloc(&x, sizeof(double) * N), "hipMalloc");
  // fill kind 4 of raja_cdna4_fill: FP64 N(0,1)
  check_lib(raja_cdna4_fill(4, x, N, 5, 0.0, 0.0, st), "fill");
  check(hipStreamSynchronize(st), "sync");

  using EXEC = RAJA::cdna4_exec<256>;
  using RED = RAJA::cdna4_reduce;
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  std::vector<double> t_old, t_par, t_nt, h_old, h_par;
  Result r_old, r_par, r_nt;

  for (int r = 0; r < reps + 1; ++r)
  {
    const bool keep = r > 0;
    float ms = 0.f;
    check(hipStreamSynchronize(st), "sync");
    {
      const auto h0 = std::chrono::steady_clock::now();
      RAJA::ReduceSum<RED, double> sum(0.0);
      RAJA::ReduceMin<RED, double> mn(std::numeric_limits<double>::max());
      RAJA::ReduceMax<RED, double> mx(-std::numeric_limits<double>::max());
      RAJA::ReduceMinLoc<RED, double> mloc;
      check(hipEventRecord(e0, st), "record");
      RAJA::forall<EXEC>(res, RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
        const double v = x[i];
        sum += v;
        mn.min(v);
        mx.max(v);
        mloc.minloc(v, i);
      });
      check(hipEventRecord(e1, st), "record");
      r_old.sum = sum.get();
      r_old.mn = mn.get();
      r_old.mx = mx.get();
      r_old.lv = mloc.get();
      r_old.loc = mloc.getLoc();
      const double hs = std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
      check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      if (keep)
      {
        t_old.push_back(ms);
        h_old.push_back(1e3 * hs);
      }
    }
    check(hipStreamSynchronize(st), "sync");
    {
      const auto h0 = std::chrono::steady_clock::now();
      double sum = 0.0, mn = std::numeric_limits<double>::max(), mx = -std::numeric_limits<double>::max();
      RAJA::expt::ValLoc<double> ml(std::numeric_limits<double>::max(), -1);
      using VSUM = RAJA::expt::ValOp<double, RAJA::operators::plus>;
      using VMIN = RAJA::expt::ValOp<double, RAJA::operators::minimum>;
      using VMAX = RAJA::expt::ValOp<double, RAJA::operators::maximum>;
      using VLMIN = RAJA::expt::ValLocOp<double, RAJA::Index_type, RAJA::operators::minimum>;
      check(hipEventRecord(e0, st), "record");
      RAJA::forall<EXEC>(res, RAJA::RangeSegment(0, N), RAJA::expt::Reduce<RAJA::operators::plus>(&sum),
                         RAJA::expt::Reduce<RAJA::operators::minimum>(&mn),
                         RAJA::expt::Reduce<RAJA::operators::maximum>(&mx),
                         RAJA::expt::Reduce<RAJA::operators::minimum>(&ml),
                         [=] RAJA_DEVICE(RAJA::Index_type i, VSUM& s, VMIN& lo, VMAX& hi, VLMIN& l) {
                           const double v = x[i];
                           s += v;
                           lo.min(v);
                           hi.max(v);
                           l.minloc(v, i);
                         });
      check(hipEventRecord(e1, st), "record");
      check(hipEventSynchronize(e1), "sync");
      const double hs = std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
      r_par.sum = sum;
      r_par.mn = mn;
      r_par.mx = mx;
      r_par.lv = ml.getVal();
      r_par.loc = ml.getLoc();
      check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      if (keep)
      {
        t_par.push_back(ms);
        h_par.push_back(1e3 * hs);
      }
    }
    check(hipStreamSynchronize(st), "sync");
    {
      // The old-style loop again, its body reading x with a nontemporal load
      // (read-once data: it bypasses the caches' allocation, as the library's
      // own reduction kernels do).  Not RAJA API: it shows what the same loop
      // reaches when the body marks its stream.  Same values, same bits.
      RAJA::ReduceSum<RED, double> sum(0.0);
      RAJA::ReduceMin<RED, double> mn(std::numeric_limits<double>::max());
      RAJA::ReduceMax<RED, double> mx(-std::numeric_limits<double>::max());
      RAJA::ReduceMinLoc<RED, double> mloc;
      check(hipEventRecord(e0, st), "record");
      RAJA::forall<EXEC>(res, RAJA::RangeSegment(0, N), [=] RAJA_DEVICE(RAJA::Index_type i) {
        const double v = __builtin_nontemporal_load(&x[i]);
        sum += v;
        mn.min(v);
        mx.max(v);
        mloc.minloc(v, i);
      });
      check(hipEventRecord(e1, st), "record");
      r_nt.sum = sum.get();
      r_nt.mn = mn.get();
      r_nt.mx = mx.get();
      r_nt.lv = mloc.get();
      r_nt.loc = mloc.getLoc();
      check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      if (keep) t_nt.push_back(ms);
    }
  }

  std::vector<double> hx(N);
  check(hipMemcpy(hx.data(), x, sizeof(double) * N, hipMemcpyDeviceToHost), "copy back");
  double s = 0.0, c = 0.0, ab = 0.0, hmn = hx[0], hmx = hx[0];
  RAJA::Index_type im = 0;
  for (RAJA::Index_type i = 0; i < N; ++i)
  {
    const double p = hx[i];
    const double t = s + p;
    c += (std::fabs(s) >= std::fabs(p)) ? (s - t) + p : (p - t) + s;
    s = t;
    ab += std::fabs(p);
    if (p < hx[im]) im = i;
    hmn = p < hmn ? p : hmn;
    hmx = hmx < p ? p : hmx;
  }
  const double kn = s + c;
  const double bound = 2.0 * log2n * std::ldexp(1.0, -53) * ab;
  auto good = [&](const Result& q) {
    return std::fabs(q.sum - kn) <= bound && q.mn == hmn && q.mx == hmx && q.lv == hx[im] && q.loc == im;
  };
  const bool same = r_old.sum == r_par.sum && r_old.mn == r_par.mn && r_old.mx == r_par.mx &&
                    r_old.lv == r_par.lv && r_old.loc == r_par.loc;
  const bool same_nt = r_old.sum == r_nt.sum && r_old.mn == r_nt.mn && r_old.mx == r_nt.mx &&
                       r_old.lv == r_nt.lv && r_old.loc == r_nt.loc;
  const bool ok = good(r_old) && good(r_par) && same && same_nt;
  const double bytes = 8.0 * N;
  auto gbs = [&](double ms) { return bytes / (ms * 1e-3) / 1e9; };
  std::printf("{\"n\": %lld, \"reps\": %d, \"pass\": %s, \"styles_bit_identical\": %s, \"sum\": %.17g, "
              "\"sum_abs_err_vs_kn\": %.3g, \"sum_bound\": %.3g, \"min\": %.17g, \"max\": %.17g, "
              "\"minloc\": [%.17g, %lld], "
              "\"reducers\": {\"ms\": %.4f, \"GBs\": %.1f, \"hbm_frac\": %.4f, \"host_ms_incl_get\": %.4f}, "
              "\"params\": {\"ms\": %.4f, \"GBs\": %.1f, \"hbm_frac\": %.4f, \"host_ms_incl_resolve\": %.4f}, "
              "\"reducers_nontemporal_body\": {\"ms\": %.4f, \"GBs\": %.1f, \"hbm_frac\": %.4f, \"same_bits\": %s}}\n",
              (long long)N, reps, ok ? "true" : "false", same ? "true" : "false", r_old.sum,
              std::fabs(r_old.sum - kn), bound, r_old.mn, r_old.mx, r_old.lv, (long long)r_old.loc,
              median(t_old), gbs(median(t_old)), gbs(median(t_old)) / 8000.0, median(h_old), median(t_par),
              gbs(median(t_par)), gbs(median(t_par)) / 8000.0, median(h_par), median(t_nt),
              gbs(median(t_nt)), gbs(median(t_nt)) / 8000.0, same_nt ? "true" : "false");
  check(hipFree(x), "free");
  return ok ? 0 : 3;
}
