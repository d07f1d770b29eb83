/*
 * raja_cdna4.h -- C ABI of libraja_cdna4.so, the MI355X (gfx950) execution
 * library behind the RAJA::cdna4_* policy family.
 *
 * RAJA has no C ABI for its algorithms: user loop bodies are compiled into
 * kernels by the C++ templates in include/RAJA/ (forall, reducers).  The
 * fixed-type primitives behind scan, sort, span reductions, segmented
 * (batched) work and request serving are exported here with plain pointers
 * and sizes, and the C++ front-end calls them for pointer + arithmetic +
 * less/greater cases -- the same envelope the reference's HIP backend accepts
 * (include/RAJA/policy/hip/sort.hpp:356-401).
 *
 * Conventions for every entry point:
 *   - returns RAJA_CDNA4_SUCCESS (0) or an error code; raja_cdna4_last_error()
 *     gives the message of the calling thread's last failure;
 *   - data pointers are caller-owned device memory (or any memory the GPU
 *     can address); temporaries come from the library's per-stream scratch;
 *   - work is enqueued asynchronously on `stream` (a hipStream_t, NULL for
 *     the null stream) unless the entry point says otherwise;
 *   - calls on different streams may run concurrently.
 */
#ifndef RAJA_CDNA4_H
#define RAJA_CDNA4_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status, types, operators ------------------------------------------- */

enum raja_cdna4_status
{
  RAJA_CDNA4_SUCCESS = 0,
  RAJA_CDNA4_ERROR_INVALID_ARGUMENT = 1,
  RAJA_CDNA4_ERROR_UNSUPPORTED = 2,
  RAJA_CDNA4_ERROR_HIP = 3,
  RAJA_CDNA4_ERROR_OUT_OF_MEMORY = 4,
  RAJA_CDNA4_ERROR_ADMISSION = 5
};

/* Element types.  Values of sort_pairs are opaque payloads of val_bytes. */
enum raja_cdna4_dtype
{
  RAJA_CDNA4_I32 = 0,
  RAJA_CDNA4_I64 = 1,
  RAJA_CDNA4_U32 = 2,
  RAJA_CDNA4_U64 = 3,
  RAJA_CDNA4_F32 = 4,
  RAJA_CDNA4_F64 = 5
};

/* Binary operators of RAJA::operators (include/RAJA/util/Operators.hpp):
 * plus (:362-377), minimum (:555-574), maximum (:576-595),
 * bit_or/bit_and (:499-527; integer types only). */
enum raja_cdna4_op
{
  RAJA_CDNA4_OP_PLUS = 0,
  RAJA_CDNA4_OP_MIN = 1,
  RAJA_CDNA4_OP_MAX = 2,
  RAJA_CDNA4_OP_BIT_OR = 3,
  RAJA_CDNA4_OP_BIT_AND = 4
};

/* ---- runtime -------------------------------------------------------------- */

/* Select `device` for the calling thread and cache its properties.
 * Replaces camp::resources::Hip device selection and the hipInfo set-up of
 * include/RAJA/policy/hip/MemUtils_HIP.hpp:166-174. */
int raja_cdna4_init(int device);

/* Text of the calling thread's last error (never NULL). */
const char* raja_cdna4_last_error(void);

/* Library version string. */
const char* raja_cdna4_version(void);

/* Compute units of the current device (256 on MI355X). */
int raja_cdna4_num_cus(void);

/* Wait for all work on `stream`.  Replaces resources::Hip::wait() /
 * synchronize_impl(hip_synchronize) (include/RAJA/policy/hip/synchronize.hpp:40-43). */
int raja_cdna4_stream_synchronize(void* stream);

/* Release the per-stream scratch the library holds for `stream` (call before
 * destroying a stream that was passed to the library). */
int raja_cdna4_stream_release(void* stream);

/* Waits for the current device, then reports whether any single-pass
 * kernel (scan or radix pass) gave up waiting for a predecessor tile's
 * lookback descriptor since the last call: such a wait is bounded (about 2.7 s
 * of wall clock) so that a broken descriptor protocol returns an error here
 * instead of hanging the device; results of the affected calls are invalid.
 * Clears the record.  No reference counterpart: the reference's scans and
 * sorts are rocPRIM calls whose errors surface through
 * CAMP_HIP_API_INVOKE_AND_CHECK (include/RAJA/policy/hip/scan.hpp:72-260). */
int raja_cdna4_device_status(void);

/* Device and pinned-host allocation through the library's pools; used by the
 * C++ front-end's resources::Cdna4 and reducer storage.  Replaces the
 * allocators and mempools of include/RAJA/policy/hip/MemUtils_HIP.hpp:69-160
 * and include/RAJA/util/basic_mempool.hpp:54-454. */
int raja_cdna4_device_malloc(void** ptr, size_t bytes);
int raja_cdna4_device_free(void* ptr);
int raja_cdna4_pinned_malloc(void** ptr, size_t bytes);
int raja_cdna4_pinned_free(void* ptr);

/* Kernel timing: when enabled, every library kernel launch is bracketed by
 * HIP events on its launch stream; times accumulate per kernel name
 * ("onesweep_pass", "radix_hist", "exclusive_scan", "reduce_sum", ...).
 * Counterpart of the reference's per-launch plugin hooks
 * (include/RAJA/util/PluginStrategy.hpp:20-40) for timing tools. */
int raja_cdna4_timing_enable(int on);
int raja_cdna4_timing_reset(void);
int raja_cdna4_timing_query(const char* kernel, int64_t* count, double* total_ms,
                            double* min_ms, double* max_ms);
/* Comma-separated names of kernels with timing records. */
int raja_cdna4_timing_names(char* buf, size_t len);

/* ---- synthetic data (portable counter-based generator) --------------------- */
/* Element i takes splitmix64((seed << 40) + i) (published splitmix64 output
 * function).  kind 0: u32 top `bits` bits (a = bits); kind 1: i32 uniform in
 * [a, b); kind 2: i64 uniform in [a, b); kind 3: f64 U[a, b); kind 4: f64
 * N(0,1) by Box-Muller on counters 2k, 2k+1; kind 5: i32 iota (a + i);
 * kind 6: f64 a + i; kind 7: u32 constant a. */
int raja_cdna4_fill(int kind, void* out, int64_t n, uint64_t seed, double a,
                    double b, void* stream);

/* ---- reductions over a contiguous span ------------------------------------ */

/* out_dev[0] = op(*init, fold_op(x[0..n))) with RAJA's identity for the
 * empty fold; min/max keep the init value on ties.  Deterministic: the
 * combine tree depends only on n.  Replaces ReduceSum/Min/Max/BitOr/BitAnd
 * over a forall (include/RAJA/policy/hip/reduce.hpp:960-1161,
 * include/RAJA/policy/hip/params/reduce.hpp:19-60) for span inputs.
 * `init` is a host pointer to one element; `out_dev` a device pointer. */
int raja_cdna4_reduce(int dtype, int op, const void* x, int64_t n,
                      const void* init, void* out_dev, void* stream);

/* MinLoc / MaxLoc (is_max = 0 / 1): value and index of the first extremum in
 * index order, combined with (init_val, init_loc), which wins ties
 * (include/RAJA/pattern/detail/reduce.hpp:127-358; ValLoc semantics of
 * include/RAJA/pattern/params/params_base.hpp:9-53).  out_val_dev holds one
 * element, out_loc_dev one int64. */
int raja_cdna4_reduce_loc(int dtype, int is_max, const void* x, int64_t n,
                          const void* init_val, int64_t init_loc,
                          void* out_val_dev, int64_t* out_loc_dev, void* stream);

/* One read of x feeding ReduceSum, ReduceMin, ReduceMax and ReduceMinLoc at
 * once (the four reducers of one forall; identities as initial values). */
typedef struct raja_cdna4_multi_f64
{
  double sum;
  double min;
  double max;
  double minloc_val;
  int64_t minloc_loc;
  int64_t pad;
} raja_cdna4_multi_f64;

int raja_cdna4_multi_reduce_f64(const double* x, int64_t n,
                                raja_cdna4_multi_f64* out_dev, void* stream);

/* ---- scans ----------------------------------------------------------------- */
/* RAJA::impl::scan::inclusive / inclusive_inplace
 * (include/RAJA/policy/hip/scan.hpp:51-160): out[0] = in[0],
 * out[i] = op(out[i-1], in[i]).  in == out is the in-place variant.
 * op is PLUS, MIN or MAX.  Integer results are bit-exact. */
int raja_cdna4_inclusive_scan(int dtype, int op, const void* in, void* out,
                              int64_t n, void* stream);

/* RAJA::impl::scan::exclusive / exclusive_inplace
 * (include/RAJA/policy/hip/scan.hpp:162-274): out[0] = *init,
 * out[i] = op(out[i-1], in[i-1]).  `init` is a host pointer to one element. */
int raja_cdna4_exclusive_scan(int dtype, int op, const void* in, void* out,
                              int64_t n, const void* init, void* stream);

/* Allocation-free raja_cdna4_exclusive_scan.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Same arguments, semantics and bits as
 * raja_cdna4_exclusive_scan, but it calls no allocator: its lookback
 * descriptors live in a fixed 1 MiB device array that each GPU holds until
 * the process ends (looked up by raja_cdna4_init, or by the first call on a
 * device).  Sizes beyond that array -- more than 131072 tiles of 32768
 * 4-byte elements, or 65536 tiles of 12288 8-byte elements (n > 805306368)
 * -- fail with RAJA_CDNA4_ERROR_UNSUPPORTED.  Every call on one GPU shares
 * the array, so calls that overlap on one GPU (concurrent streams, or
 * threads) race and give wrong results: serialise them. */
int raja_cdna4_exclusive_scan_noalloc(int dtype, int op, const void* in, void* out,
                                      int64_t n, const void* init, void* stream);

/* ---- sorts ------------------------------------------------------------------- */
/* RAJA::impl::sort::{stable,unstable}_pairs (include/RAJA/policy/hip/sort.hpp:
 * 417-495, 653-663): sort keys ascending (descending = 0, operators::less) or
 * descending (1, operators::greater); values follow their keys.  The LSD
 * radix sort is stable, so both variants give the stable result.  vals may be
 * NULL (keys only: RAJA::impl::sort::{stable,unstable}, sort.hpp:77-356).
 * val_bytes is 4 or 8.  Signed and floating keys order as operator< does
 * (-0.0 and +0.0 are equal keys and keep their input order; NaN keys have no
 * defined place).  Keys with every digit pass constant are not
 * re-sorted for that pass. */
int raja_cdna4_sort_pairs(int key_dtype, void* keys, void* vals, int val_bytes,
                          int64_t n, int descending, void* stream);

/* ---- shape tuning ------------------------------------------------------------- */
/* The reference chooses launch shapes at run time with its occupancy
 * calculator and concretizers (include/RAJA/policy/hip/MemUtils_HIP.hpp:
 * 389-519, 547-633) and at compile time with tuning policies
 * (include/RAJA/policy/hip/policy.hpp:1228-1353).  Here the tile of a span
 * scan (vectors per lane) and of a radix pass (keys per thread, 4-byte keys)
 * are measured per (operation, element type, length bucket ceil(log2 n)) on
 * the current device and kept in a text file named for the device's
 * architecture, CU count and HIP runtime version, in $RAJA_CDNA4_TUNE_DIR
 * (else $XDG_CACHE_HOME/raja_cdna4, else $HOME/.cache/raja_cdna4).  Scans and
 * sorts look their bucket up on every call and take the default shape when
 * it is absent; with RAJA_CDNA4_AUTOTUNE=1 in the environment an absent
 * bucket is measured at its first call instead (never under graph capture).
 * Integer scans and sorts give the same bits in every shape, so the file is
 * safe to delete: only speed depends on it. */
enum raja_cdna4_tune_op
{
  RAJA_CDNA4_TUNE_INCLUSIVE_SCAN = 0,
  RAJA_CDNA4_TUNE_EXCLUSIVE_SCAN = 1,
  RAJA_CDNA4_TUNE_SORT_PAIRS = 2
};

/* Measure every candidate shape of (tune_op, dtype) on n synthetic elements
 * (allocated for the measurement) on `stream`, record the fastest for n's
 * bucket and rewrite the cache file.  Synchronous. */
int raja_cdna4_autotune(int tune_op, int dtype, int64_t n, void* stream);

/* The shape a call of (tune_op, dtype, n) takes: *choice 0 is the default
 * shape, else vectors per lane (scans) or keys per thread (sorts); *tuned is
 * 1 when the bucket has a recorded measurement. */
int raja_cdna4_tune_query(int tune_op, int dtype, int64_t n, int* choice, int* tuned);

/* Path of the current device's cache file ("" when no directory is known). */
int raja_cdna4_tune_cache_path(char* buf, size_t len);

/* Forget every recorded shape in this process; the file is read again at
 * the next call. */
int raja_cdna4_tune_forget(void);

/* ---- batched (segmented) work ---------------------------------------------- */
/* No single reference entry point: these replace one reference algorithm call
 * per range -- a forall with ReduceSum, an exclusive_scan, a sort_pairs
 * (include/RAJA/policy/hip/{reduce,scan,sort}.hpp) -- issued for each of many
 * small uneven ranges, with one or two kernels over all of them.
 *
 * A segment plan is built once from host offsets (nseg + 1 non-decreasing
 * element offsets; segment s is [offsets[s], offsets[s+1])) and holds the
 * device segment/tile tables; it is reusable across calls and streams and
 * its calls can be captured in a hipGraph. */
typedef struct raja_cdna4_segments_s* raja_cdna4_segments;

int raja_cdna4_segments_create(const int64_t* offsets, int64_t nseg, raja_cdna4_segments* out);
int raja_cdna4_segments_destroy(raja_cdna4_segments plan);

/* out_dev[s] = op(*init, fold(x[segment s])) for every segment, with the same
 * combine tree as raja_cdna4_reduce on that segment alone (bit-identical
 * results).  dtype I32, I64, F32 or F64; op PLUS, MIN or MAX. */
int raja_cdna4_segmented_reduce(raja_cdna4_segments plan, int dtype, int op, const void* x,
                                const void* init, void* out_dev, void* stream);

/* Exclusive / inclusive scan of every segment independently (each starts
 * from *init / its first element); in == out allowed. */
int raja_cdna4_segmented_exclusive_scan(raja_cdna4_segments plan, int dtype, int op,
                                        const void* in, void* out, const void* init,
                                        void* stream);
int raja_cdna4_segmented_inclusive_scan(raja_cdna4_segments plan, int dtype, int op,
                                        const void* in, void* out, void* stream);

/* Stable sort of every segment's (key, value) pairs independently. */
int raja_cdna4_segmented_sort_pairs(raja_cdna4_segments plan, int key_dtype, void* keys,
                                    void* vals, int val_bytes, int descending, void* stream);

/* ---- request serving ------------------------------------------------------
 *
 * A server runs a stream of independent requests (sum, exclusive scan,
 * sort_pairs), each on caller-owned device inputs and outputs, with every
 * byte of working memory carved from one device arena sized at creation.
 * Requests are admitted in arrival order while their working memory fits in
 * the arena and fewer than max_inflight launch units are running; a request
 * whose working memory exceeds the whole arena is rejected with
 * RAJA_CDNA4_ERROR_ADMISSION.  In batched mode, small requests that are
 * waiting when a unit slot frees up are coalesced into one batch (one table
 * upload and one or two kernels per request kind); large ones run alone;
 * units go round-robin over the server's streams.  In per-request mode each
 * request is its own unit on one stream: the reference's one algorithm call
 * per request (a RAJA::forall with ReduceSum, RAJA::exclusive_scan or
 * RAJA::stable_sort_pairs, include/RAJA/pattern/forall.hpp:545-627,
 * include/RAJA/pattern/scan.hpp:216-295, include/RAJA/pattern/sort.hpp:219-301,
 * each allocating its temporaries from a pool, include/RAJA/util/basic_mempool.hpp:54-454).
 * The RAJA counterpart of coalescing is WorkGroup (many small loops in one
 * launch, include/RAJA/pattern/WorkGroup.hpp:117-530).
 *
 * Every request's result is bit-identical whether it runs alone, in a batch
 * or in either mode (sums use the canonical reduction tree; scans and sorts
 * are exact).  Request inputs must not alias any output of a request that
 * may be in flight at the same time, and such requests' outputs must not
 * overlap: requests in flight run in no fixed order. */

enum raja_cdna4_request_kind
{
  RAJA_CDNA4_REQ_SUM_F64 = 0,           /* out[0] = sum of in[0..n), FP64          */
  RAJA_CDNA4_REQ_EXCLUSIVE_SCAN_I32 = 1, /* out = exclusive plus-scan of in, init 0 */
  RAJA_CDNA4_REQ_SORT_PAIRS_U32_I32 = 2  /* stable ascending sort of (in, in_vals)
                                            into (out, out_vals)                     */
};

enum raja_cdna4_serve_mode
{
  RAJA_CDNA4_SERVE_BATCHED = 0,
  RAJA_CDNA4_SERVE_PER_REQUEST = 1
};

typedef struct
{
  int mode;            /* raja_cdna4_serve_mode                                     */
  int streams;         /* streams for launch units, 1..16 (per-request mode uses 1)  */
  int max_inflight;    /* launch units in flight, 1..64                              */
  int max_batch;       /* requests per batch, >= 1                                   */
  int64_t batch_limit; /* requests with n <= batch_limit join batches (sorts of up
                          to 8192 pairs always); at most 2^27 (the batched sum
                          folds at most 4096 tiles of 32768 doubles)                 */
  int64_t arena_bytes; /* device working memory of the server                       */
  int honor_arrivals;  /* 1: a request starts no earlier than its arrival_s;
                          0: every request has arrived at time 0 (saturation)        */
  int pad;
  int64_t batch_bytes; /* a batch stops growing once the algorithmic bytes of its
                          requests (8n sum, 8n scan, 68n sort) reach this; 0: no
                          cap.  Caps how long one unit runs, so that requests
                          arriving behind it wait less                             */
} raja_cdna4_serve_config;

typedef struct
{
  int kind;           /* raja_cdna4_request_kind                                     */
  int pad;
  int64_t n;          /* elements (pairs for sorts)                                  */
  const void* in;     /* FP64 / int32 / uint32 keys                                  */
  const void* in_vals;/* int32 values (sorts)                                        */
  void* out;          /* 1 double / n int32 / n uint32 keys                          */
  void* out_vals;     /* n int32 values (sorts)                                      */
  double arrival_s;   /* seconds after the run starts; non-decreasing                */
} raja_cdna4_request;

typedef struct
{
  int status;         /* RAJA_CDNA4_SUCCESS or RAJA_CDNA4_ERROR_ADMISSION            */
  int unit;           /* index of the launch unit the request ran in (-1: none)       */
  int unit_size;      /* requests in that unit                                        */
  int pad;
  double start_s;     /* launched, seconds after the run start                        */
  double done_s;      /* completion seen by the server, seconds after the run start   */
} raja_cdna4_result;

typedef struct
{
  int64_t ticket;           /* from raja_cdna4_server_submit                        */
  raja_cdna4_result result; /* times in seconds after raja_cdna4_server_start       */
} raja_cdna4_completion;

typedef struct raja_cdna4_server_s* raja_cdna4_server;

/* Creates the server's streams and arena on the current device. */
int raja_cdna4_server_create(const raja_cdna4_serve_config* config, raja_cdna4_server* server);
/* Runs nreq requests to completion (blocking) and fills results[0..nreq);
 * *wall_s (optional) is the time from the run start to the last completion.
 * A replay of a known request list (arrival_s honoured when the config asks);
 * refused while the server is started. */
int raja_cdna4_server_run(raja_cdna4_server server, const raja_cdna4_request* requests,
                          int64_t nreq, raja_cdna4_result* results, double* wall_s);

/* Open serving: start the server's scheduling thread, which admits
 * requests as they are submitted (batching those waiting together, exactly
 * as server_run does) while earlier units run. */
int raja_cdna4_server_start(raja_cdna4_server server);
/* Queue n requests (copied; their buffers must stay valid until their
 * completion is polled); tickets[i] (optional) receives request i's ticket.
 * arrival_s is ignored: a request arrives when it is submitted.  Safe to call
 * from several threads. */
int raja_cdna4_server_submit(raja_cdna4_server server, const raja_cdna4_request* requests,
                             int64_t n, int64_t* tickets);
/* Moves up to max completions into out[0..*count), waiting up to timeout_s
 * for the first one when none is ready and requests are outstanding. */
int raja_cdna4_server_poll(raja_cdna4_server server, raja_cdna4_completion* out, int64_t max,
                           double timeout_s, int64_t* count);
/* Waits until every submitted request has completed (their completions stay
 * to be polled), then ends the scheduling thread.  destroy stops first. */
int raja_cdna4_server_stop(raja_cdna4_server server);
int raja_cdna4_server_destroy(raja_cdna4_server server);

#ifdef __cplusplus
}
#endif

#endif /* RAJA_CDNA4_H */
