/*
 * raja_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of RAJA's seq_exec and omp_parallel_for_exec semantics for
 * the data-parallel hot path (forall + reducers, scans, sorts, WorkGroup halo
 * pack/unpack, segmented reductions/scans).  It is the checker the GPU
 * results are compared against, and it is the "port" CPU baseline that
 * bench.py times.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Parity pinning (what each group is checked against, beyond itself):
 *   rng      : splitmix64 published first output for state 0, and an
 *              independent Python restatement (tests/test_oracle.py).
 *   reduce   : closed forms (iota sums), numpy argmin/argmax first-index
 *              rule, integer-valued FP data (exact), Kahan-Neumaier vs
 *              math.fsum (correctly rounded).
 *   scan     : numpy.cumsum / minimum.accumulate / maximum.accumulate and
 *              the triangular-number closed form of the reference's scan
 *              tests (iota inputs, N in {0,357,32000}).
 *   sort     : numpy.argsort(kind="stable") (a library stable sort).
 *   segmented: per-segment numpy reductions/scans.
 *   halo     : the example's closed-form extents (list lengths, first and
 *              last index of every list) and the pack->unpack round trip.
 *   omp_*    : equal to the seq versions bit-for-bit for integer data.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Portable counter-based RNG: splitmix64 of (seed << 40) + i.               */
/* ------------------------------------------------------------------------ */

static inline uint64_t orc_mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

ORC_API uint64_t orc_rng_raw(uint64_t seed, uint64_t i)
{
  return orc_mix64((seed << 40) + i + 0x9E3779B97F4A7C15ull);
}

/* keys uniform over [0, 2^bits), 1 <= bits <= 32 */
ORC_API void orc_fill_u32_bits(uint32_t* out, int64_t n, uint64_t seed, int bits)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = (uint32_t)(orc_rng_raw(seed, (uint64_t)i) >> (64 - bits));
}

/* uniform integers in [lo, hi) by the multiply-shift map of the top 32 bits */
ORC_API void orc_fill_i32_range(int32_t* out, int64_t n, uint64_t seed,
                                int32_t lo, int32_t hi)
{
  uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = (int32_t)((int64_t)lo +
                       (int64_t)(((orc_rng_raw(seed, (uint64_t)i) >> 32) * span) >> 32));
}

ORC_API void orc_fill_i64_range(int64_t* out, int64_t n, uint64_t seed,
                                int64_t lo, int64_t hi)
{
  uint64_t span = (uint64_t)(hi - lo);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = lo + (int64_t)(((orc_rng_raw(seed, (uint64_t)i) >> 32) * span) >> 32);
}

/* U[a,b): a + (b-a) * k * 2^-53 with k the top 53 bits */
ORC_API void orc_fill_f64_uniform(double* out, int64_t n, uint64_t seed,
                                  double a, double b)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double u = (double)(orc_rng_raw(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
    out[i] = a + (b - a) * u;
  }
}

/* N(0,1) by Box-Muller on counters 2k, 2k+1; element 2k is r*cos, 2k+1 is r*sin */
ORC_API void orc_fill_f64_normal(double* out, int64_t n, uint64_t seed)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = (uint64_t)i >> 1;
    double u1 = ((double)(orc_rng_raw(seed, 2 * k) >> 11) + 1.0) * 0x1.0p-53;
    double u2 = (double)(orc_rng_raw(seed, 2 * k + 1) >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    double t = 6.283185307179586 * u2;
    out[i] = (i & 1) ? r * sin(t) : r * cos(t);
  }
}

/* ------------------------------------------------------------------------ */
/* Reducers: include/RAJA/pattern/detail/reduce.hpp:282-593 (seq combinable) */
/* The loop body's copy starts at the identity and folds contributions left  */
/* to right; its destructor folds it into the parent (init) only when it      */
/* differs from the identity (:316-331).  min/max keep lhs on ties           */
/* (include/RAJA/util/Operators.hpp:555-595).                                 */
/* ------------------------------------------------------------------------ */

ORC_API double orc_seq_sum_f64(const double* x, int64_t n, double init)
{
  double child = 0.0;
  for (int64_t i = 0; i < n; ++i) child += x[i];
  return (child != 0.0) ? init + child : init;
}

ORC_API double orc_seq_dot_f64(const double* x, const double* y, int64_t n, double init)
{
  double child = 0.0;
  for (int64_t i = 0; i < n; ++i) child += x[i] * y[i];
  return (child != 0.0) ? init + child : init;
}

/* Kahan-Neumaier compensated sum plus sum of |x| (the FP64 error-bound oracle). */
ORC_API void orc_kn_sum_f64(const double* x, int64_t n, double* sum, double* abssum)
{
  double s = 0.0, c = 0.0, a = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double v = x[i];
    double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else                    c += (v - t) + s;
    s = t;
    a += fabs(v);
  }
  *sum = s + c;
  *abssum = a;
}

/* dot product x.y with compensation on the exact product error (TwoProduct via fma) */
ORC_API void orc_kn_dot_f64(const double* x, const double* y, int64_t n,
                            double* sum, double* abssum)
{
  double s = 0.0, c = 0.0, a = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double p = x[i] * y[i];
    double pe = fma(x[i], y[i], -p);
    double t = s + p;
    if (fabs(s) >= fabs(p)) c += (s - t) + p;
    else                    c += (p - t) + s;
    c += pe;
    s = t;
    a += fabs(p);
  }
  *sum = s + c;
  *abssum = a;
}

#define ORC_SEQ_MINMAX(T, SFX)                                                \
  ORC_API T orc_seq_min_##SFX(const T* x, int64_t n, T init, T ident)         \
  {                                                                            \
    T child = ident;                                                           \
    for (int64_t i = 0; i < n; ++i) child = (x[i] < child) ? x[i] : child;    \
    if (child != ident) init = (child < init) ? child : init;                  \
    return init;                                                               \
  }                                                                            \
  ORC_API T orc_seq_max_##SFX(const T* x, int64_t n, T init, T ident)         \
  {                                                                            \
    T child = ident;                                                           \
    for (int64_t i = 0; i < n; ++i) child = (child < x[i]) ? x[i] : child;    \
    if (child != ident) init = (init < child) ? child : init;                  \
    return init;                                                               \
  }                                                                            \
  /* ValueLoc compares values only (detail/reduce.hpp:170-178): first hit wins */ \
  ORC_API void orc_seq_minloc_##SFX(const T* x, int64_t n, T init, int64_t init_loc, \
                                    T ident, T* out_val, int64_t* out_loc)     \
  {                                                                            \
    T cv = ident; int64_t cl = -1;                                             \
    for (int64_t i = 0; i < n; ++i) if (x[i] < cv) { cv = x[i]; cl = i; }     \
    if (cv != ident && cv < init) { init = cv; init_loc = cl; }                \
    *out_val = init; *out_loc = init_loc;                                      \
  }                                                                            \
  ORC_API void orc_seq_maxloc_##SFX(const T* x, int64_t n, T init, int64_t init_loc, \
                                    T ident, T* out_val, int64_t* out_loc)     \
  {                                                                            \
    T cv = ident; int64_t cl = -1;                                             \
    for (int64_t i = 0; i < n; ++i) if (cv < x[i]) { cv = x[i]; cl = i; }     \
    if (cv != ident && init < cv) { init = cv; init_loc = cl; }                \
    *out_val = init; *out_loc = init_loc;                                      \
  }

ORC_SEQ_MINMAX(double, f64)
ORC_SEQ_MINMAX(float, f32)
ORC_SEQ_MINMAX(int32_t, i32)
ORC_SEQ_MINMAX(int64_t, i64)
ORC_SEQ_MINMAX(uint32_t, u32)
ORC_SEQ_MINMAX(uint64_t, u64)

#define ORC_SEQ_INTSUM(T, SFX)                                                 \
  ORC_API T orc_seq_sum_##SFX(const T* x, int64_t n, T init)                   \
  {                                                                            \
    T child = 0;                                                               \
    for (int64_t i = 0; i < n; ++i) child = (T)(child + x[i]);                 \
    return (child != 0) ? (T)(init + child) : init;                            \
  }                                                                            \
  ORC_API T orc_seq_bitor_##SFX(const T* x, int64_t n, T init)                 \
  {                                                                            \
    T child = 0;                                                               \
    for (int64_t i = 0; i < n; ++i) child |= x[i];                             \
    return (child != 0) ? (T)(init | child) : init;                            \
  }                                                                            \
  ORC_API T orc_seq_bitand_##SFX(const T* x, int64_t n, T init)                \
  {                                                                            \
    T child = (T)~(T)0;                                                        \
    for (int64_t i = 0; i < n; ++i) child &= x[i];                             \
    return (child != (T)~(T)0) ? (T)(init & child) : init;                     \
  }

ORC_SEQ_INTSUM(int32_t, i32)
ORC_SEQ_INTSUM(int64_t, i64)
ORC_SEQ_INTSUM(uint32_t, u32)
ORC_SEQ_INTSUM(uint64_t, u64)

ORC_API float orc_seq_sum_f32(const float* x, int64_t n, float init)
{
  float child = 0.0f;
  for (int64_t i = 0; i < n; ++i) child += x[i];
  return (child != 0.0f) ? init + child : init;
}

/* ------------------------------------------------------------------------ */
/* Scans: include/RAJA/policy/sequential/scan.hpp:43-153.  op: 0 plus,       */
/* 1 minimum, 2 maximum.  Inclusive: out[0]=in[0], out[i]=op(out[i-1],in[i]). */
/* Exclusive: out[0]=init, out[i]=op(out[i-1], in[i-1]); written in the      */
/* in-place-safe form of exclusive_inplace (:79-97) so in==out is allowed.   */
/* ------------------------------------------------------------------------ */

#define ORC_OP(T, op, a, b)                                                    \
  ((op) == 0 ? (T)((a) + (b)) : (op) == 1 ? (((b) < (a)) ? (b) : (a))          \
                              : (((a) < (b)) ? (b) : (a)))

#define ORC_SEQ_SCAN(T, SFX)                                                   \
  ORC_API void orc_seq_inclusive_##SFX(const T* in, T* out, int64_t n, int op) \
  {                                                                            \
    if (n <= 0) return;                                                        \
    T agg = in[0];                                                             \
    out[0] = agg;                                                              \
    for (int64_t i = 1; i < n; ++i) { agg = ORC_OP(T, op, agg, in[i]); out[i] = agg; } \
  }                                                                            \
  ORC_API void orc_seq_exclusive_##SFX(const T* in, T* out, int64_t n, int op, T init) \
  {                                                                            \
    T agg = init;                                                              \
    for (int64_t i = 0; i < n; ++i) { T t = in[i]; out[i] = agg; agg = ORC_OP(T, op, agg, t); } \
  }

ORC_SEQ_SCAN(int32_t, i32)
ORC_SEQ_SCAN(int64_t, i64)
ORC_SEQ_SCAN(uint32_t, u32)
ORC_SEQ_SCAN(uint64_t, u64)
ORC_SEQ_SCAN(float, f32)
ORC_SEQ_SCAN(double, f64)

/* ------------------------------------------------------------------------ */
/* Sorts: stable sort of (key, value) pairs by key, the semantics of RAJA's  */
/* merge_sort with compare_first (include/RAJA/util/sort.hpp:576-693,        */
/* include/RAJA/policy/sequential/sort.hpp:125-156).  desc=0 -> less,        */
/* desc=1 -> greater.  Values are opaque 32- or 64-bit payloads (vbytes).    */
/* Bottom-up merge with a scratch buffer; ties take the left run first.      */
/* ------------------------------------------------------------------------ */

#define ORC_SORT(T, SFX)                                                       \
  static void orc_merge_##SFX(const T* ka, const unsigned char* va, T* kb,     \
                              unsigned char* vb, int64_t lo, int64_t mid,      \
                              int64_t hi, int desc, int vbytes)                \
  {                                                                            \
    int64_t i = lo, j = mid, o = lo;                                           \
    while (i < mid && j < hi) {                                                \
      int take_right = desc ? (ka[i] < ka[j]) : (ka[j] < ka[i]);               \
      int64_t s = take_right ? j++ : i++;                                      \
      kb[o] = ka[s];                                                           \
      if (vbytes) memcpy(vb + o * vbytes, va + s * vbytes, (size_t)vbytes);    \
      ++o;                                                                     \
    }                                                                          \
    while (i < mid) { kb[o] = ka[i]; if (vbytes) memcpy(vb + o * vbytes, va + i * vbytes, (size_t)vbytes); ++o; ++i; } \
    while (j < hi)  { kb[o] = ka[j]; if (vbytes) memcpy(vb + o * vbytes, va + j * vbytes, (size_t)vbytes); ++o; ++j; } \
  }                                                                            \
  static void orc_msort_##SFX(T* keys, void* vals, int64_t n, int desc, int vbytes, \
                              int64_t first_width)                             \
  {                                                                            \
    if (n <= 1) return;                                                        \
    T* kt = (T*)malloc((size_t)n * sizeof(T));                                 \
    unsigned char* vt = vbytes ? (unsigned char*)malloc((size_t)n * (size_t)vbytes) : 0; \
    T* ka = keys; unsigned char* va = (unsigned char*)vals;                    \
    T* kb = kt;   unsigned char* vb = vt;                                      \
    for (int64_t w = first_width; w < n; w *= 2) {                             \
      for (int64_t lo = 0; lo < n; lo += 2 * w) {                              \
        int64_t mid = lo + w < n ? lo + w : n;                                 \
        int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;                          \
        orc_merge_##SFX(ka, va, kb, vb, lo, mid, hi, desc, vbytes);            \
      }                                                                        \
      T* tk = ka; ka = kb; kb = tk;                                            \
      unsigned char* tv = va; va = vb; vb = tv;                                \
    }                                                                          \
    if (ka != keys) {                                                          \
      memcpy(keys, ka, (size_t)n * sizeof(T));                                 \
      if (vbytes) memcpy(vals, va, (size_t)n * (size_t)vbytes);                \
    }                                                                          \
    free(kt); free(vt);                                                        \
  }                                                                            \
  static void orc_isort_##SFX(T* keys, unsigned char* vals, int64_t n, int desc, int vbytes) \
  {                                                                            \
    unsigned char tmpv[8];                                                     \
    for (int64_t i = 1; i < n; ++i) {                                          \
      T k = keys[i];                                                           \
      if (vbytes) memcpy(tmpv, vals + i * vbytes, (size_t)vbytes);             \
      int64_t j = i;                                                           \
      while (j > 0 && (desc ? (keys[j - 1] < k) : (k < keys[j - 1]))) {        \
        keys[j] = keys[j - 1];                                                 \
        if (vbytes) memcpy(vals + j * vbytes, vals + (j - 1) * vbytes, (size_t)vbytes); \
        --j;                                                                   \
      }                                                                        \
      keys[j] = k;                                                             \
      if (vbytes) memcpy(vals + j * vbytes, tmpv, (size_t)vbytes);             \
    }                                                                          \
  }                                                                            \
  ORC_API void orc_seq_stable_sort_pairs_##SFX(T* keys, void* vals, int64_t n, \
                                               int desc, int vbytes)           \
  {                                                                            \
    if (n <= 1) return;                                                        \
    const int64_t run = 16;                                                    \
    for (int64_t lo = 0; lo < n; lo += run)                                    \
      orc_isort_##SFX(keys + lo, vbytes ? (unsigned char*)vals + lo * vbytes : 0, \
                      (n - lo < run) ? n - lo : run, desc, vbytes);            \
    orc_msort_##SFX(keys, vals, n, desc, vbytes, run);                         \
  }                                                                            \
  ORC_API void orc_seq_stable_sort_##SFX(T* keys, int64_t n, int desc)        \
  {                                                                            \
    orc_seq_stable_sort_pairs_##SFX(keys, 0, n, desc, 0);                      \
  }                                                                            \
  /* omp_parallel_for_exec stable sort: per-thread ranges by firstIndex, then */ \
  /* a pairwise merge tree (include/RAJA/policy/openmp/sort.hpp:100-138).    */ \
  ORC_API void orc_omp_stable_sort_pairs_##SFX(T* keys, void* vals, int64_t n, \
                                               int desc, int vbytes, int nthreads) \
  {                                                                            \
    if (n <= 1) return;                                                        \
    if (nthreads < 1) nthreads = 1;                                            \
    T* kt = (T*)malloc((size_t)n * sizeof(T));                                 \
    unsigned char* vt = vbytes ? (unsigned char*)malloc((size_t)n * (size_t)vbytes) : 0; \
    int p = nthreads;                                                          \
    _Pragma("omp parallel for num_threads(p) schedule(static,1)")              \
    for (int t = 0; t < p; ++t) {                                              \
      int64_t b = (int64_t)(((uint64_t)n * (uint64_t)t) / (uint64_t)p);        \
      int64_t e = (int64_t)(((uint64_t)n * (uint64_t)(t + 1)) / (uint64_t)p);  \
      orc_seq_stable_sort_pairs_##SFX(keys + b, vbytes ? (unsigned char*)vals + b * vbytes : 0, \
                                      e - b, desc, vbytes);                    \
    }                                                                          \
    T* ka = keys; unsigned char* va = (unsigned char*)vals; T* kb = kt; unsigned char* vb = vt; \
    for (int m = 1; m < p; m *= 2) {                                           \
      _Pragma("omp parallel for num_threads(p) schedule(dynamic,1)")           \
      for (int t = 0; t < p; t += 2 * m) {                                     \
        int tm = (t + m < p) ? t + m : p, te = (t + 2 * m < p) ? t + 2 * m : p; \
        int64_t lo = (int64_t)(((uint64_t)n * (uint64_t)t) / (uint64_t)p);    \
        int64_t mid = (int64_t)(((uint64_t)n * (uint64_t)tm) / (uint64_t)p);  \
        int64_t hi = (int64_t)(((uint64_t)n * (uint64_t)te) / (uint64_t)p);   \
        orc_merge_##SFX(ka, va, kb, vb, lo, mid, hi, desc, vbytes);            \
      }                                                                        \
      T* tk = ka; ka = kb; kb = tk; unsigned char* tv = va; va = vb; vb = tv;  \
    }                                                                          \
    if (ka != keys) {                                                          \
      memcpy(keys, ka, (size_t)n * sizeof(T));                                 \
      if (vbytes) memcpy(vals, va, (size_t)n * (size_t)vbytes);                \
    }                                                                          \
    free(kt); free(vt);                                                        \
  }

ORC_SORT(uint32_t, u32)
ORC_SORT(int32_t, i32)
ORC_SORT(float, f32)
ORC_SORT(uint64_t, u64)
ORC_SORT(int64_t, i64)
ORC_SORT(double, f64)

/* ------------------------------------------------------------------------ */
/* omp_parallel_for_exec baselines (include/RAJA/policy/openmp/ reducer, scan, sort headers) */
/* ------------------------------------------------------------------------ */

ORC_API int orc_omp_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* One forall over x with ReduceSum/Min/Max/MinLoc: per-thread copies fold
 * into the parent under a critical section (openmp/reduce.hpp:46-69). */
ORC_API void orc_omp_multi_reduce_f64(const double* x, int64_t n, int nthreads,
                                      double* sum, double* mn, double* mx,
                                      double* mlv, int64_t* mll)
{
  double S = 0.0, MN = 1.7976931348623157e308, MX = -1.7976931348623157e308;
  double LV = 1.7976931348623157e308; int64_t LL = -1;
#pragma omp parallel num_threads(nthreads)
  {
    double s = 0.0, a = 1.7976931348623157e308, b = -1.7976931348623157e308;
    double lv = 1.7976931348623157e308; int64_t ll = -1;
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < n; ++i) {
      double v = x[i];
      s += v;
      a = (v < a) ? v : a;
      b = (b < v) ? v : b;
      if (v < lv) { lv = v; ll = i; }
    }
#pragma omp critical
    {
      S += s;
      MN = (a < MN) ? a : MN;
      MX = (MX < b) ? b : MX;
      if (lv < LV || (lv == LV && ll < LL)) { LV = lv; LL = ll; }
    }
  }
  *sum = S; *mn = MN; *mx = MX; *mlv = LV; *mll = LL;
}

/* ReduceSum<omp_reduce, double> over one forall: each thread's copy starts at
 * 0 and folds its static chunk in index order; the copies fold into the
 * parent under a critical section as they are destroyed
 * (include/RAJA/policy/openmp/reduce.hpp:46-69), so the bits depend on
 * thread arrival order. */
ORC_API double orc_omp_sum_f64(const double* x, int64_t n, int nthreads)
{
  double S = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    double s = 0.0;
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < n; ++i) s += x[i];
#pragma omp critical
    S += s;
  }
  return S;
}

/* daxpy y = a*x + y then dot(x,y) and minloc(x) in separate foralls */
ORC_API void orc_omp_daxpy_dot_minloc(const double* x, double* y, int64_t n,
                                      double a, int nthreads, double* dot,
                                      double* mlv, int64_t* mll)
{
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) y[i] = a * x[i] + y[i];
  double D = 0.0, LV = 1.7976931348623157e308; int64_t LL = -1;
#pragma omp parallel num_threads(nthreads)
  {
    double d = 0.0, lv = 1.7976931348623157e308; int64_t ll = -1;
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < n; ++i) {
      d += x[i] * y[i];
      if (x[i] < lv) { lv = x[i]; ll = i; }
    }
#pragma omp critical
    {
      D += d;
      if (lv < LV || (lv == LV && ll < LL)) { LV = lv; LL = ll; }
    }
  }
  *dot = D; *mlv = LV; *mll = LL;
}


/* Two-level scans of include/RAJA/policy/openmp/scan.hpp:46-181: each thread
 * scans its firstIndex chunk (include/RAJA/pattern/detail/algorithm.hpp:60-65)
 * sequentially, one thread scans the chunk totals, then every thread adds
 * its carry.  Plus only (the configs' op). */
#define ORC_OMP_SCAN(T, SFX)                                                   \
  ORC_API void orc_omp_inclusive_##SFX(const T* in, T* out, int64_t n, int nthreads) \
  {                                                                            \
    if (n <= 0) return;                                                        \
    int p = (int)((n < nthreads) ? n : nthreads);                              \
    T* sums = (T*)calloc((size_t)p + 1, sizeof(T));                            \
    _Pragma("omp parallel num_threads(p)")                                     \
    {                                                                          \
      int t = omp_get_thread_num(), pp = omp_get_num_threads();               \
      int64_t b = (int64_t)(((uint64_t)n * (uint64_t)t) / (uint64_t)pp);      \
      int64_t e = (int64_t)(((uint64_t)n * (uint64_t)(t + 1)) / (uint64_t)pp); \
      T agg = 0;                                                               \
      for (int64_t i = b; i < e; ++i) { agg = (T)(agg + in[i]); out[i] = agg; } \
      sums[t + 1] = agg;                                                       \
      _Pragma("omp barrier")                                                   \
      _Pragma("omp single")                                                    \
      for (int k = 1; k <= pp; ++k) sums[k] = (T)(sums[k] + sums[k - 1]);      \
      T carry = sums[t];                                                       \
      if (carry != 0) for (int64_t i = b; i < e; ++i) out[i] = (T)(out[i] + carry); \
    }                                                                          \
    free(sums);                                                                \
  }                                                                            \
  ORC_API void orc_omp_exclusive_##SFX(const T* in, T* out, int64_t n, T init, int nthreads) \
  {                                                                            \
    if (n <= 0) return;                                                        \
    int p = (int)((n < nthreads) ? n : nthreads);                              \
    T* sums = (T*)calloc((size_t)p + 1, sizeof(T));                            \
    _Pragma("omp parallel num_threads(p)")                                     \
    {                                                                          \
      int t = omp_get_thread_num(), pp = omp_get_num_threads();               \
      int64_t b = (int64_t)(((uint64_t)n * (uint64_t)t) / (uint64_t)pp);      \
      int64_t e = (int64_t)(((uint64_t)n * (uint64_t)(t + 1)) / (uint64_t)pp); \
      T agg = (t == 0) ? init : (T)0;                                          \
      for (int64_t i = b; i < e; ++i) { T v = in[i]; out[i] = agg; agg = (T)(agg + v); } \
      sums[t + 1] = agg;                                                       \
      _Pragma("omp barrier")                                                   \
      _Pragma("omp single")                                                    \
      for (int k = 1; k <= pp; ++k) sums[k] = (T)(sums[k] + sums[k - 1]);      \
      T carry = sums[t];                                                       \
      if (carry != 0) for (int64_t i = b; i < e; ++i) out[i] = (T)(out[i] + carry); \
    }                                                                          \
    free(sums);                                                                \
  }

ORC_OMP_SCAN(int32_t, i32)
ORC_OMP_SCAN(int64_t, i64)

/* ------------------------------------------------------------------------ */
/* Segmented work (per-segment seq calls).  offsets has nseg+1 entries.      */
/* ------------------------------------------------------------------------ */

ORC_API void orc_seg_sum_f64(const double* x, const int64_t* offsets, int64_t nseg,
                             double* out)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < nseg; ++s)
    out[s] = orc_seq_sum_f64(x + offsets[s], offsets[s + 1] - offsets[s], 0.0);
}

ORC_API void orc_seg_kn_sum_f64(const double* x, const int64_t* offsets, int64_t nseg,
                                double* out, double* abssum)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < nseg; ++s)
    orc_kn_sum_f64(x + offsets[s], offsets[s + 1] - offsets[s], &out[s], &abssum[s]);
}

ORC_API void orc_seg_exclusive_i32(const int32_t* in, int32_t* out,
                                   const int64_t* offsets, int64_t nseg, int32_t init)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < nseg; ++s)
    orc_seq_exclusive_i32(in + offsets[s], out + offsets[s],
                          offsets[s + 1] - offsets[s], 0, init);
}

ORC_API void orc_seg_inclusive_i32(const int32_t* in, int32_t* out,
                                   const int64_t* offsets, int64_t nseg)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < nseg; ++s)
    orc_seq_inclusive_i32(in + offsets[s], out + offsets[s],
                          offsets[s + 1] - offsets[s], 0);
}

ORC_API void orc_seg_stable_sort_pairs_u32(uint32_t* keys, void* vals,
                                           const int64_t* offsets, int64_t nseg,
                                           int desc, int vbytes)
{
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t s = 0; s < nseg; ++s)
    orc_seq_stable_sort_pairs_u32(keys + offsets[s],
                                  (unsigned char*)vals + offsets[s] * vbytes,
                                  offsets[s + 1] - offsets[s], desc, vbytes);
}

/* ------------------------------------------------------------------------ */
/* Halo exchange index lists (examples/tut_halo-exchange.cpp:1793-2062).     */
/* 26 neighbours in the example's order; per axis, direction -1/0/+1 maps to */
/* pack range [h,2h) / [h,g+h) / [g,g+h) and unpack range [0,h) / [h,g+h) /  */
/* [g+h,g+2h).  Lists enumerate k outermost, then j, then i.                 */
/* ------------------------------------------------------------------------ */

static const signed char orc_halo_dirs[26][3] = {
  {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},
  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},
  {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1},
  {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
  {1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {1, 1, 1}};

static void orc_halo_range(int dir, int g, int h, int unpack, int* lo, int* hi)
{
  if (dir == 0) { *lo = h; *hi = g + h; }
  else if (dir < 0) { *lo = unpack ? 0 : h; *hi = unpack ? h : 2 * h; }
  else { *lo = unpack ? g + h : g; *hi = unpack ? g + 2 * h : g + h; }
}

/* lengths[26]; when lists != NULL it receives the concatenated lists, list l
 * starting at the prefix sum of lengths.  Returns the total length. */
ORC_API int64_t orc_halo_lists(const int* grid_dims, int h, int unpack,
                               int32_t* lengths, int32_t* lists)
{
  const int sj = grid_dims[0] + 2 * h;
  const int sk = sj * (grid_dims[1] + 2 * h);
  int64_t pos = 0;
  for (int l = 0; l < 26; ++l) {
    int r[3][2];
    for (int a = 0; a < 3; ++a)
      orc_halo_range(orc_halo_dirs[l][a], grid_dims[a], h, unpack, &r[a][0], &r[a][1]);
    int32_t len = (r[0][1] - r[0][0]) * (r[1][1] - r[1][0]) * (r[2][1] - r[2][0]);
    lengths[l] = len;
    if (lists) {
      for (int kk = r[2][0]; kk < r[2][1]; ++kk)
        for (int jj = r[1][0]; jj < r[1][1]; ++jj)
          for (int ii = r[0][0]; ii < r[0][1]; ++ii)
            lists[pos++] = ii + jj * sj + kk * sk;
    } else {
      pos += len;
    }
  }
  return pos;
}

/* One seq halo cycle over nvars variables (tut_halo-exchange.cpp:298-340):
 * pack buffer l holds var0[list], var1[list], ... contiguously; then unpack.
 * buffers: total_len * nvars doubles laid out list by list. */
ORC_API void orc_halo_cycle(double* const* vars, int nvars,
                            const int32_t* pack_lists, const int32_t* unpack_lists,
                            const int32_t* lengths, double* buffers)
{
  int64_t off = 0;
  for (int l = 0; l < 26; ++l) {
    double* buf = buffers + off * nvars;
    const int32_t* list = pack_lists + off;
    for (int v = 0; v < nvars; ++v)
      for (int32_t i = 0; i < lengths[l]; ++i) buf[(int64_t)v * lengths[l] + i] = vars[v][list[i]];
    off += lengths[l];
  }
  off = 0;
  for (int l = 0; l < 26; ++l) {
    const double* buf = buffers + off * nvars;
    const int32_t* list = unpack_lists + off;
    for (int v = 0; v < nvars; ++v)
      for (int32_t i = 0; i < lengths[l]; ++i) vars[v][list[i]] = buf[(int64_t)v * lengths[l] + i];
    off += lengths[l];
  }
}

/* The example's omp_parallel_for_exec variant (tut_halo-exchange.cpp:615-700):
 * one parallel forall per (neighbour, variable) loop, pack loops then unpack
 * loops, same buffer layout as orc_halo_cycle. */
ORC_API void orc_omp_halo_cycle(double* const* vars, int nvars,
                                const int32_t* pack_lists, const int32_t* unpack_lists,
                                const int32_t* lengths, double* buffers, int nthreads)
{
  int64_t off = 0;
  for (int l = 0; l < 26; ++l) {
    double* buf = buffers + off * nvars;
    const int32_t* list = pack_lists + off;
    const int32_t len = lengths[l];
    for (int v = 0; v < nvars; ++v) {
      double* b = buf + (int64_t)v * len;
      const double* var = vars[v];
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int32_t i = 0; i < len; ++i) b[i] = var[list[i]];
    }
    off += len;
  }
  off = 0;
  for (int l = 0; l < 26; ++l) {
    const double* buf = buffers + off * nvars;
    const int32_t* list = unpack_lists + off;
    const int32_t len = lengths[l];
    for (int v = 0; v < nvars; ++v) {
      const double* b = buf + (int64_t)v * len;
      double* var = vars[v];
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int32_t i = 0; i < len; ++i) var[list[i]] = b[i];
    }
    off += len;
  }
}
