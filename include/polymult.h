/* polymult.h -- C ABI of the MI355X negacyclic multiplier.
 *
 *   c = a * b mod (X^N + 1, q),  a, b, c: N coefficients of L little-endian
 *   u32 limbs (row-major [N][L]), q a W-bit modulus of L limbs.
 *
 * Drop-in boundary. The reference binds three C++ entry points
 * (fherma.h:46-48, bound by name in fherma.toml:7-10):
 *   void* fherma_init(const fherma::Point&)            -> pm_init
 *   fherma::Outputs fherma_run(void*, const Inputs&)   -> pm_run
 *   void fherma_free(void*)                            -> pm_free
 * Those three keep their C++ signatures in fherma.h (this directory) and are
 * thin wrappers over the functions below, which take plain pointers and sizes
 * so that any FFI (ctypes, cgo, JNI) can bind them.
 *
 * Errors: every int-returning call returns 0 on success and -1 on failure;
 * pm_last_error() then returns the message (per thread), worded as the
 * reference's exceptions (solve.cu:873-913, 1482-1483).
 * Threading: one state is used by one thread at a time (the reference's
 * state is not re-entrant either); distinct states are independent.
 */
#ifndef POLYMULT_H
#define POLYMULT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_state pm_state;

typedef struct pm_info {
    uint32_t N, W, L;
    uint32_t K;        /* RNS primes */
    uint64_t delta;    /* 2^W - q */
    uint32_t w1, w2;   /* NTT split: 2^w1 rows x 2^w2 columns */
    uint64_t device_bytes;  /* device memory the state holds */
} pm_info;

/* Validates the point, builds the tables on the device, warms up.
 * Replaces fherma_init (fherma.h:46; solve.cu:1392-1478). q has q_limbs
 * limbs, which must equal L. */
int pm_init(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs, pm_state** out);

/* c = a * b from and to host memory (pageable or pinned), N*L u32 each,
 * blocking. Replaces fherma_run (fherma.h:47; solve.cu:1480-1529). */
int pm_run(pm_state* s, const uint32_t* a, const uint32_t* b, uint32_t* c);

/* c = a * b with a, b, c already in device memory (N*L u32 each). Returns
 * once enqueued. With a `stream` (a hipStream_t) the call runs in order on
 * it. With NULL, consecutive calls rotate over the state's internal lanes
 * (three by default, PM_LANES=1..4; each with its own residue buffers), so one
 * call's kernels overlap the next call's; results still land in call order
 * (a call's final kernel waits for the previous call's), and pm_sync waits
 * for every lane. Calls on different streams (and pm_run, the sharded call)
 * share the state's default residue buffers: each waits on the device for
 * the one before, so they stay correct but do not overlap. Inputs
 * must not change before pm_sync. No counterpart in the reference: this is
 * the device-resident form of the same call. */
int pm_run_device(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, void* stream);

/* The same multiply as pm_run_device, with no allocation on the call: its two
 * residue arrays live in a fixed array in the library's device code (96 MiB
 * per GPU, held from the library's load to the end of the process), so the
 * call makes no hipMalloc, hipFree or other allocation, and no event wait or
 * record. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). Limit: 2 K N <= 25,165,824 residue words (K N <=
 * 12,582,912; N = 65536 with up to 192 primes, W up to about 2,950 bits);
 * larger points fail with a message. Runs in order on `stream` (NULL: the
 * state's own stream); it is not ordered against the state's calls on other
 * streams or lanes (pm_sync, or the caller's stream, orders them). Every call
 * on one GPU, from any state, uses that GPU's one array: calls that overlap on
 * a GPU race. Nothing in the array needs clearing between calls (each call
 * writes every residue it reads). */
int pm_run_device_scratch(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, void* stream);

/* Releases everything. Replaces fherma_free (fherma.h:48; solve.cu:1531-1535). */
void pm_free(pm_state* s);

/* The calling thread's last error message, worded as the reference's
 * exceptions (solve.cu:873-913, 1482-1483), which the C ABI cannot throw. */
const char* pm_last_error(void);
/* The plan's sizes. No counterpart in the reference (its state is opaque). */
int pm_query(const pm_state* s, pm_info* out);

/* Host-only point check (no GPU needed): validates like pm_init and reports
 * K, delta and up to `cap` primes, as the reference's precompute
 * (solve.cu:869-1020) derives them. */
int pm_point_info(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs, uint32_t* K,
                  uint64_t* delta, uint32_t* primes, size_t cap);

/* ---- device buffers and timing, for callers without another GPU runtime ----
 * No counterpart in the reference, whose only entry points take host
 * buffers; these serve the device-resident calls above. */
int pm_dev_alloc(size_t bytes, void** out);
int pm_dev_free(void* p);
int pm_copy_to_device(void* dst, const void* src, size_t bytes);
int pm_copy_to_host(void* dst, const void* src, size_t bytes);
/* A hipStream_t (non-blocking, default priority) to pass as `stream` to the
 * device entry points, its host-side wait, and its release. */
int pm_stream_create(void** out);
int pm_stream_sync(void* stream);
int pm_stream_destroy(void* stream);
/* Waits for every device call of the state (every lane). No counterpart in
 * the reference (its fherma_run returns when c is on the host). */
int pm_sync(pm_state* s);

/* Average milliseconds, from HIP events on the state's stream, of `iters`
 * launches enqueued back to back (after one warm launch): ms[1] RNS, ms[2]
 * column forward, ms[3] fused rows, ms[4] column inverse, ms[5] CRT (each
 * kernel repeated alone, on the residues the chain leaves), then ms[0] the
 * whole multiply. d_c holds a * b afterwards. No counterpart in the
 * reference: a timing diagnostic of the device-resident form. */
int pm_time_device(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int iters, float* ms);

/* PCIe and host-copy rates (GB/s, best of `reps`) for transfers of `bytes`:
 * out[0] pinned H2D, out[1] pinned D2H, out[2] pageable H2D, out[3] pageable
 * D2H, out[4] host staging copy on 1 thread, out[5] on `threads` threads,
 * out[6] H2D + D2H of bytes/2 each concurrently. Diagnostic; the reference's
 * counterpart is its PM_DIAG setup benchmark (solve.cu:1245-1358). */
int pm_probe_transfers(size_t bytes, int threads, int reps, double* out);

/* ---- prime shards: the multi-GPU split, shown on one GPU ----
 * No counterpart in the reference (single GPU). The K primes split into
 * `shards` ranges of whole groups of 8 primes (1 <= shards <= ceil(K/8));
 * bounds[0..shards] receives them. A shard runs RNS, the NTT passes and the
 * linear part of the CRT (int64 slot partials, (L + 3) x N) for its primes
 * only, as one GPU of a sharded multiply would; the combine sums the partials
 * (where GPUs would reduce-scatter them) and finishes the CRT.
 * pm_run_device_sharded runs the shards one after another on this GPU and
 * gives the same c as pm_run_device. pm_time_shard times, averaged over
 * `iters`: ms[0] the whole sharded multiply here, ms[1..5] shard 0's RNS,
 * forward columns, fused rows, inverse columns and CRT partial, ms[6] the
 * combine. */
int pm_shard_bounds(pm_state* s, int shards, uint32_t* bounds);
int pm_run_device_sharded(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int shards,
                          void* stream);
int pm_time_shard(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int shards, int iters,
                  float* ms);

/* HBM copy rate (GB/s, read + write) of a plain device copy kernel over
 * `bytes` (a multiple of 16), averaged over `iters` launches. Diagnostic: the
 * reference point for the HBM traffic counters and for achievable bandwidth.
 * No counterpart in the reference. */
int pm_probe_copy(size_t bytes, int iters, double* gbps);

/* Issue rates of four VALU instructions, in tera lane-operations per
 * second: out[0] v_add_u32, out[1] v_mul_lo_u32, out[2] v_mul_hi_u32,
 * out[3] v_mad_u64_u32 (8 waves per SIMD, 8 independent chains per lane,
 * `iters` x 128 instructions per lane). Diagnostic: the measured peak the NTT
 * kernels' modular multiplies are priced against. No counterpart in the
 * reference. */
int pm_probe_valu(int iters, double* out);

/* ---- stage entry points, for stage-by-stage parity tests ----
 * All on device buffers, synchronous.
 *   rns:     R[K][N] = a mod p_j
 *   ntt_fwd: R <- forward negacyclic NTT of each prime row (in place),
 *            bit-reversed output order
 *   ntt_inv: R <- N * inverse NTT (unscaled)
 *   crt:     c[N][L] from canonical residues X[K][N] of the integer product
 * The reference exposes no such entry points; each stage computes what one
 * of its kernels computes: rns = k_decompose (solve.cu:147-202), ntt_fwd =
 * k_col<false> then the forward rows (solve.cu:254-355), ntt_inv = the inverse
 * rows then k_col<true> (solve.cu:254-284, 357-378), crt = k_crt
 * (solve.cu:380-501). */
int pm_stage_rns(pm_state* s, const uint32_t* d_a, uint32_t* d_R);
int pm_stage_ntt_fwd(pm_state* s, uint32_t* d_R);
int pm_stage_ntt_inv(pm_state* s, uint32_t* d_R);
int pm_stage_crt(pm_state* s, const uint32_t* d_X, uint32_t* d_c);

/* ---- batched form: many products of uneven N and W in one graph launch ----
 * No counterpart in the reference (it multiplies one pair per call). Products
 * share one RNS prime set (p = 1 mod 2^17) and one twiddle table; their K
 * follows the same rule as pm_init. q_concat holds each product's q (L_i =
 * ceil(W_i / 32) limbs) back to back. One batch runs at a time. */
typedef struct pm_batch pm_batch;
int pm_batch_create(uint32_t count, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat, pm_batch** out);
void pm_batch_free(pm_batch* b);
const char* pm_batch_last_error(void);
int pm_batch_product_info(const pm_batch* b, uint32_t i, pm_info* out);
/* Enqueues every product (device buffers, N_i*L_i u32 each, 16-byte aligned)
 * as one hipGraph launch on `stream` (NULL: the batch's own): five phase
 * launches, each over all products' tiles. */
int pm_batch_run(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                 void* stream);
/* Waits for the device (and so for every launch of the batch). */
int pm_batch_sync(pm_batch* b);
/* Average milliseconds of one batch launch over `iters` graph replays. */
int pm_batch_time(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                  int iters, float* ms);
/* Sets the launch configuration: column tile 8/16/32, ceil_split 0/1 (odd
 * logN: the larger half on the column pass), blocks_per_cu 0 (one block per
 * tile) or 1..8 (persistent blocks claiming tiles), group_mb (products run in
 * groups of at most that many MiB of residues; 0: one group; a group of one
 * product runs the single-product kernels), branches 1..3 (graph chains the
 * groups go round-robin over). */
int pm_batch_configure(pm_batch* b, int col_tile, int ceil_split, int blocks_per_cu, int group_mb, int branches);
/* Times every candidate (column tile 8/16/32, 4-step split of odd logN,
 * tiles handed out by the dispatcher (0) or by 2 persistent blocks per CU,
 * product groups of 48/96/192 MiB of residues or one group, one to three
 * graph branches), keeps the fastest; best[0..5] = column tile, split, blocks per
 * CU, microseconds per launch, group MiB, branches. */
int pm_batch_autotune(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                      int iters, int* best);

/* ---- serving: requests from host buffers, continuous batching ----
 * No counterpart in the reference (its harness calls fherma_run once per
 * case). A server holds a catalog of points (shared prime set, as the batched
 * form). Requests name a point and pass pageable host buffers (N*L u32 each;
 * the caller keeps them alive until pm_server_wait reports the request done).
 * A scheduler thread groups the waiting requests into one batch whenever a
 * batch slot is free, admitting them in arrival order while their device
 * footprint (2 N L + 2 K N words) and pinned staging (2 N L words) fit the
 * budgets; a request larger than a budget is refused at submission. */
typedef struct pm_server pm_server;
typedef struct {
    uint64_t device_budget; /* HBM bytes for requests in flight (0: 4 GiB) */
    uint64_t pinned_budget; /* pinned staging bytes (0: 1 GiB) */
    uint32_t max_batch;     /* requests per batch (0: 64) */
    uint32_t max_inflight;  /* batches in flight, at most 8 (0: 3) */
    uint32_t host_threads;  /* copy threads, half staging, half scattering (0: min(8, cpus - 2)) */
    int32_t col_tile;       /* column tile 8, 16 or 32 (0: 32) */
    uint64_t max_batch_bytes; /* staged input bytes per batch, above one request (0: 64 MiB) */
} pm_server_config;
typedef struct {
    uint64_t submitted, completed, failed, batches;
    uint64_t deferred; /* times the queue head waited for memory */
    double mean_batch, p50_ms, p99_ms, max_ms, mean_ms;
    /* per batch, mean ms: host staging, issuing copies and launches, issue
     * to results back in pinned memory, scatter to the callers' buffers */
    double stage_ms, issue_ms, gpu_ms, scatter_ms;
    double sched_busy, completer_busy; /* fractions of wall time since reset */
    uint64_t device_budget, device_peak, pinned_budget, pinned_peak;
} pm_server_stats;
typedef struct {
    double offered_rate, seconds;
    uint64_t requests, completed, failed;
    double achieved_rate, p50_ms, p99_ms, max_ms, mean_ms, mean_batch;
    uint64_t batches, deferred, device_peak;
    double max_late_ms; /* worst delay of a submission behind its arrival time */
    double stage_ms, issue_ms, gpu_ms, scatter_ms, sched_busy, completer_busy; /* as pm_server_stats */
    /* latency split at admission into a batch: queue = arrival to admission,
     * service = admission to result written */
    double queue_p50_ms, queue_p99_ms, service_p50_ms, service_p99_ms;
    double point_p50_ms[8], point_p99_ms[8]; /* per catalog point (first 8) */
    double cpu_util;           /* process CPU seconds per wall second during the run */
    uint64_t ctx_invol;        /* involuntary context switches of the process */
    uint64_t driver_ctx_invol; /* ... of the driving thread */
    /* results compared bit for bit with the plain path's product of the same
     * operands (pm_init / pm_run), and how many differed */
    uint64_t verified, mismatches;
} pm_serve_report;
int pm_server_create(uint32_t npoints, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat,
                     const pm_server_config* cfg, pm_server** out);
/* Finishes every submitted request, then stops. */
void pm_server_free(pm_server* s);
const char* pm_server_last_error(void);
int pm_server_submit(pm_server* s, uint32_t point, const uint32_t* a, const uint32_t* b, uint32_t* c,
                     uint64_t* ticket);
/* 0: done (the ticket is consumed; latency from submission to c written),
 * 1: not done within timeout_s, -1: error (the request failed, or a bad
 * ticket). */
int pm_server_wait(pm_server* s, uint64_t ticket, double timeout_s, double* latency_ms);
/* Counters and latency percentiles since the last reset. */
int pm_server_stats_get(pm_server* s, int reset, pm_server_stats* out);
/* Open-loop Poisson load from the calling thread: arrivals at `rate` per
 * second for `seconds`, point i drawn with weight weights[i], operands
 * uniform in [0, q) generated from `seed` (two pairs per point, alternating);
 * latency counts from each scheduled arrival. Before the clock starts, the
 * plain path (pm_init / pm_run) computes the product of every pair; each
 * completed request's result is compared with it (report: verified,
 * mismatches). */
int pm_server_drive(pm_server* s, double rate, double seconds, const double* weights, uint64_t seed,
                    pm_serve_report* out);
/* Closed-loop load: `depth` requests kept in flight for `seconds`, a new one
 * submitted as each completes; achieved_rate is the server's capacity on
 * the mix (offered_rate is reported as 0). */
int pm_server_saturate(pm_server* s, uint32_t depth, double seconds, const double* weights, uint64_t seed,
                       pm_serve_report* out);

#ifdef __cplusplus
}
#endif

#endif
