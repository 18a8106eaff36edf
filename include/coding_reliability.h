/*
 * coding_reliability.h -- C ABI of the MI355X distance-distribution engine.
 *
 * Plain C types only (no HIP, no torch types).  Every entry point returns a
 * cr_status and never exits the process; cr_last_error() gives the message of
 * the last failing call on the calling thread.  Histograms are "raw" unless a
 * function says otherwise: raw[b] counts pairs at Hamming distance b, as the
 * reference kernels accumulate them before host post-processing.
 *
 * Which reference interface each entry point replaces is stated next to it
 * (paths relative to the reference repository root).  The C++ entry points
 * with the reference's exact signatures live in cr_reference_api.h.
 */
#ifndef CODING_RELIABILITY_H_
#define CODING_RELIABILITY_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cr_status {
    CR_OK = 0,
    CR_ERR_INVALID_ARGUMENT = 1, /* bad size, width, or pointer */
    CR_ERR_UNSUPPORTED = 2,      /* valid request outside what this library supports */
    CR_ERR_NO_DEVICE = 3,        /* no gfx950 device at the given index */
    CR_ERR_OUT_OF_MEMORY = 4,    /* device allocation failed */
    CR_ERR_DEVICE = 5,           /* a HIP runtime call failed */
    CR_ERR_ADMISSION = 6,        /* serving: request exceeds the HBM budget */
    CR_ERR_CANCELLED = 7,        /* serving: request was cancelled */
    CR_ERR_NOT_FOUND = 8,        /* serving: unknown request id */
    CR_ERR_BUSY = 9              /* serving: result not ready yet */
} cr_status;

typedef struct cr_context cr_context;

/* Message for a status code / of the last failing call on this thread. */
const char *cr_status_string(cr_status s);
const char *cr_last_error(void);

/* Library version string and the device-code target it was built for. */
const char *cr_version(void);

/* --- context ------------------------------------------------------------
 * One context drives one device with its own HIP stream and device buffers.
 * A context is not reentrant: calls on one context must not overlap.
 * Replaces the per-call device set-up/teardown of run_ancoding
 * (distance_distribution/src/an_coding.cu:196-252,321-326): no device reset. */
cr_status cr_context_create(int device, cr_context **out);
cr_status cr_context_destroy(cr_context *ctx);

typedef struct cr_device_info {
    char name[64];
    char arch[32];
    int compute_units;
    int max_clock_mhz;
    uint64_t hbm_bytes;
} cr_device_info;
cr_status cr_device_query(cr_context *ctx, cr_device_info *out);

/* Free and total device memory (bytes) of the context's device, as the HIP
 * runtime reports them; for sizing server budgets and leak checks.  No
 * counterpart in the reference, which allocated per call. */
cr_status cr_device_memory(cr_context *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Device ceilings the histogram kernels are measured against (no
 * counterpart in the reference, which times its kernels only).  Runs one
 * probe kernel on every resident single-wave workgroup for `iters` loop
 * iterations and returns the rate it sustained:
 *   probe 0, 4: VALU lane-instructions/s of xor + popcount-accumulate on 8 or
 *               16 independent chains (the pair, grid and exhaustive Hamming
 *               inner loops);
 *   probe 1:    VALU lane-instructions/s of Philox4x32-10 (the Monte Carlo
 *               generator: 20 64-bit multiplies and 20 three-input xors per
 *               block);
 *   probe 3:    the same for Philox blocks whose words go through the AN
 *               Monte Carlo pair work (multiply, xor, popcount);
 *   probe 2, 5: lane increments/s of wave-wide ds_add_u32 to lane-private LDS
 *               counters in the histogram layout of every kernel, the lanes
 *               of an instruction in different rows (2) or one row (5);
 *   probe 6, 7: bytes/s of streaming 16-byte stores (6) or loads (7), `iters`
 *               passes (at most 64) over a 256 MiB buffer, the size of the
 *               spectral path's batches;
 *   probe 8:    lane-reads/s of ds_read_u8 from 128-byte LDS tables (one
 *               dword per bank), the Hamming kernels' syndrome lookups.
 * kernel_ms (may be NULL) receives the probe's device time. */
cr_status cr_probe_ceiling(cr_context *ctx, uint32_t probe, uint32_t iters, double *units_per_s, double *kernel_ms);

/* --- exhaustive AN ---------------------------------------------------------
 * raw[b] = #{ 0 <= i < j < num_words : popcount(A*i XOR A*j) == b } for
 * b < nbins, computed exactly on the GPU (uint64 counts).
 * Requires A >= 1, num_words >= 1 and (num_words-1)*A < 2^64; nbins must cover
 * every possible distance (bit length of (num_words-1)*A, plus one).
 * With num_words = 2^k this is the histogram accumulated by dancoding_shared
 * (distance_distribution/src/an_coding.cu:50-105) before run_ancoding doubles it
 * (an_coding.cu:287-292).  With A = A0*s and num_words = m it is the
 * integer-stride grid {A0*(i*s) : i < m}.
 * For num_words = 2^k, 14 <= k <= 32, the same histogram is computed from the code's
 * Walsh-Hadamard spectrum (cr_an_pairs_spectral) when that is estimated at
 * under 80 % of the pair kernel's time (multipliers well below 2^k);
 * cr_an_pairs_tuned always runs the pair kernel.
 * kernel_ms (may be NULL) receives the device time of the launches. */
cr_status cr_an_pairs(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                      double *kernel_ms);

/* Tuning knobs for the pair kernel; 0 = automatic: 8 rows per lane from 2^16
 * words (4 from 2^15, else 2) and column chunks giving about 16 items per
 * resident wave (the settings measured fastest for single codes; batches are
 * tuned by cr_autotune_batch). */
typedef struct cr_pair_tuning {
    uint32_t rows_per_lane;   /* codewords held in registers per lane: 2, 4, 8 or 16 */
    uint32_t chunk_blocks;    /* column chunk in units of the 64*rows row block (single code) */
    uint32_t waves;           /* persistent waves launched (0 = fill the device) */
    uint32_t items_per_wave;  /* batch: target work items per resident wave (0 = 16) */
} cr_pair_tuning;
cr_status cr_an_pairs_tuned(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                            const cr_pair_tuning *tuning, double *kernel_ms);

/* Allocation-free variant of cr_an_pairs_tuned (automatic tuning).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The pair kernel's job descriptor, work-queue head and
 * histogram copies live in a file-scope device array of 32 KiB (one per GPU,
 * held until the process ends, its address looked up once per device by
 * cr_context_create or by the first call); the call allocates nothing.  It
 * enqueues its kernels on `stream` (a hipStream_t; NULL = the null stream of
 * `device`, which it makes current) and returns without waiting: raw_device
 * (nbins uint64 in device memory) holds the raw histogram of cr_an_pairs once
 * the stream reaches that point.  Every call on a GPU uses that GPU's array
 * and starts by clearing it, so calls that overlap on one GPU race: the
 * caller orders them (one stream, or events).  The histogram must fit the
 * array: at most 65 bins (codewords of up to 64 bits). */
cr_status cr_an_pairs_noalloc(int device, uint64_t A, uint64_t num_words, uint64_t *raw_device, uint32_t nbins,
                              void *stream);

/* Same raw[] as cr_an_pairs with num_words = 2^k, computed from the
 * Walsh-Hadamard spectrum of the code instead of its pairs: for the odd part
 * A' of A (the distances of A*i and A'*i agree) and n = bit length of
 * A'*(2^k-1), one length-2^k transform per s < 2^(n-k), O(2^n * k) work
 * against the 2^(2k-1) pairs of dancoding_shared
 * (distance_distribution/src/an_coding.cu:50-105), exact integer arithmetic.
 * Supported: 14 <= k <= 32 and n <= min(2k, 48) (CR_ERR_UNSUPPORTED
 * otherwise); large k gives exact distributions where the pairs are out of
 * reach (A = 61: k = 28 in 65 ms, k = 32 in 1.03 s).
 * kernel_ms (may be NULL) receives the device time of all the launches. */
cr_status cr_an_pairs_spectral(cr_context *ctx, uint64_t A, uint32_t k, uint64_t *raw, uint32_t nbins,
                               double *kernel_ms);

/* --- reference grid approximation ------------------------------------------
 * Kernel-level histogram of run_ancoding_grid
 * (distance_distribution/src/an_coding_grid.cu:215-322, one device): after the
 * reference normalisation (A a power of two -> 1; m1, m2 clamped to 2^n), the
 * sample list v_k = A*(T)(k * 2^n/m1) produced by the reference kernel loop
 * (1D: an_coding_grid.cu:60-107 incl. its unrolled overshoot; 2D: :121-140),
 * T = uint32 or uint64 as the reference dispatches, is compared with
 *   gdim 1: every code word A*i, i < 2^n (ordered pairs, self pairs included)
 *   gdim 2: the grid words A*(T)(i * 2^n/m2), i < m2 (bin 0 not counted).
 * hist[b] gets the raw pair count at distance b (b < nbins; nbins >= 65 is
 * always enough).  Scale with cr_an_grid_report.  m1 >= 3 for gdim 1. */
cr_status cr_an_grid_ref(cr_context *ctx, int gdim, uint32_t n, uint64_t m1, uint64_t m2, uint64_t A, uint64_t *hist,
                         uint32_t nbins, double *kernel_ms);

/* Host scaling of run_ancoding_grid (an_coding_grid.cu:316-322):
 * out[i] = (uint128)((long double)factor * hist[i]), factor = 2^n/m1 (gdim 1) or
 * 4^n/(m1*m2) (gdim 2), with m1, m2 clamped as above.  out holds count_counts
 * little-endian 128-bit values as (lo, hi) uint64 pairs. */
cr_status cr_an_grid_report(int gdim, uint32_t n, uint64_t m1, uint64_t m2, const uint64_t *hist,
                            uint32_t count_counts, uint64_t *out_u128);

/* --- Extended Hamming ---------------------------------------------------------
 * Weight histogram of the non-extended codewords (popcount of data plus the
 * parity bits popcount(d & mask_t) & 1, computeHamming<N>,
 * distance_distribution/src/hamming.cu:39-59) over a stream of data words.
 * The masks are the reference's (hamming.h:54-64), extended to 64-bit data by
 * the patterns of legacy/Hamming/ExtHamming64.hpp:31-37, truncated to k bits:
 * k=8 -> (13,8), 16 -> (22,16), 24 -> (30,24), 32 -> (39,32), 64 -> (72,64). */
typedef enum cr_ham_source {
    CR_HAM_RANGE = 0,  /* d = first + s, s < count (exhaustive when first=0, count=2^k) */
    CR_HAM_STRIDE = 1, /* d = (first + s*stride) mod 2^k */
    CR_HAM_LIST = 2,   /* d = list[s] (host array of count words) */
    CR_HAM_MC = 3      /* d = Monte Carlo word (first + s) of the Philox stream (seed, stream) */
} cr_ham_source;

typedef struct cr_ham_request {
    uint32_t k;         /* data bits, 1..64 */
    uint32_t source;    /* cr_ham_source */
    uint64_t first;
    uint64_t count;
    uint64_t stride;
    const uint64_t *list;
    uint64_t seed;
    uint32_t stream;
    uint32_t reserved;
} cr_ham_request;

/* Number of parity bits (without the overall parity bit) for k data bits. */
uint32_t cr_hamming_parity_bits(uint32_t k);

/* hist[w] = samples whose non-extended codeword has weight w, w <= k + parity
 * bits (nbins must exceed that).  Monte Carlo words: sample x of the stream is
 * word x%4 (k <= 32; masked to k bits) or the word pair 2(x%2), 2(x%2)+1
 * (k > 32) of Philox4x32-10 block x/4 (x/2) with counter
 * (block, block >> 32, stream, 0x43520001) and key (seed, seed >> 32);
 * `first` must be a multiple of 4 (k <= 32) or 2.  Replaces dhamming,
 * dhamming_grid_1d and dhamming_mc (hamming.cu:64-81, hamming_grid.cu:53-74,
 * hamming_mc.cu:57-81). */
cr_status cr_hamming_weights(cr_context *ctx, const cr_ham_request *req, uint64_t *hist, uint32_t nbins,
                             double *kernel_ms);

/* n requests (1..4096) in one call: hists[i * nbins + w] is request i's
 * histogram, equal to cr_hamming_weights of that request; nbins must exceed
 * every request's largest weight.  The requests of k <= 32 run as the jobs of
 * one kernel launch, those of k > 32 as the jobs of a second; kernel_ms spans
 * both.  No counterpart in the reference: it puts small requests (the (39,32)
 * stride grid of 2^20 words, a few microseconds of work) into the launch of a
 * larger one (its 2^34-word Monte Carlo run) instead of paying a launch each. */
cr_status cr_hamming_weights_batch(cr_context *ctx, const cr_ham_request *reqs, uint32_t n, uint64_t *hists,
                                   uint32_t nbins, double *kernel_ms);

/* Post-processing of run_hamming (hamming.cu:182-208) from a non-extended
 * weight histogram: out[0] = 2^k, out[1] = N*2^k with spheres, odd weights
 * merged into the next even bin, the 1-bit-sphere closed form for odd bins,
 * bins >= 3 shifted by k.  N = k + parity bits + 1; out has N+1 uint128 values
 * as (lo, hi) pairs.  CR_ERR_UNSUPPORTED when a count exceeds 128 bits (the
 * reference's uint128_t; for such k cr_hamming_sdc_table keeps 256 bits). */
cr_status cr_hamming_report(uint32_t k, const uint64_t *weights, int with_1bit, uint64_t *out_u128);

/* SDC table of the (N,k) Extended Hamming code, k = 1..64, as printed by the
 * legacy table program (legacy/hamming_count_undetectable_errors.cpp:226-339,
 * its countHammingUndetectableErrors<k> output): per weight i = 0..N the
 * non-extended weight distribution, the detecting-only and correcting Extended
 * Hamming distributions, their totals x 2^k, 2^k * C(N, i) and the two
 * probabilities.  method 0: device enumeration of all 2^k words for k <= 40,
 * else the MacWilliams identity over the 2^r dual words (exact, any k); 1:
 * device enumeration (k <= 40); 2: MacWilliams (ctx may be NULL).  If
 * weights_u128 is not NULL it receives the N+1 non-extended weights as (lo, hi)
 * pairs (nweights >= N+1).  *text is released with cr_free. */
cr_status cr_hamming_sdc_table(cr_context *ctx, uint32_t k, int method, uint64_t *weights_u128, uint32_t nweights,
                               char **text);

/* --- AN Monte Carlo -------------------------------------------------------------
 * hist[b] = #{(i, j) : i < 2^n, j < m, popcount(A*i XOR A*u_ij) == b} with u_ij
 * sample j of row i: Philox4x32-10 block b of counter
 * (b, b >> 32, i mod 2^32, 0x43520001 ^ (i >> 32)), key (seed, seed >> 32);
 * for n <= 32, b = j/4 and u_ij = word j%4; for n > 32, b = j/2 and u_ij =
 * word 2(j%2) | word 2(j%2)+1 << 32; masked to n bits.  Replaces
 * dancoding_mc_shared's cuRAND Sobol draw (an_coding_mc.cu:94-158), which is
 * not reproducible; 1 <= n <= 63 with A*(2^n - 1) < 2^64 (the reference's
 * ccmax <= 64, an_coding_mc.cu:163-188). */
cr_status cr_an_mc(cr_context *ctx, uint64_t A, uint32_t n, uint64_t m, uint64_t seed, uint64_t *hist, uint32_t nbins,
                   double *kernel_ms);

/* AN Monte Carlo histograms of consecutive sample ranges in one launch:
 * bounds[0] <= bounds[1] <= ... <= bounds[count] (1 <= count <= 64), and
 * hists[i * nbins + b] counts the pairs (i', j) of cr_an_mc (same A, n, seed)
 * at distance b with bounds[i] <= j < bounds[i+1].  Sample j is the same draw
 * in every call with this seed, so cr_an_mc with m samples equals the sum of
 * the ranges below m, and each sample is drawn once however many ranges there
 * are.  With bounds {0, m, 3m, 9m, ...} the cumulative sums are the histograms
 * of ancoding_mc_search's successive run_ancoding_mc calls
 * (distance_distribution/src/main.cpp:179-202), which otherwise redraw every
 * earlier sample at each step. */
cr_status cr_an_mc_ranges(cr_context *ctx, uint64_t A, uint32_t n, const uint64_t *bounds, uint32_t count,
                          uint64_t seed, uint64_t *hists, uint32_t nbins, double *kernel_ms);

/* --- batched exhaustive AN ---------------------------------------------------------
 * Many codes in one persistent launch.  Code c is (A_c, num_words_c); its raw
 * histogram (as cr_an_pairs) is written to raw[c*stride .. c*stride+stride).
 * Work items of all codes share one device queue ordered longest code first,
 * so codes of very different size (k = 8 .. 24) balance across the device.
 * Replaces the one-launch-per-code loops of the super-A searches
 * (distance_distribution_super_a/src/main.cpp:194-297, which call
 * super_a/src/an_coding.cu:135-244 once per A). */
typedef struct cr_an_code {
    uint64_t A;
    uint64_t num_words;
} cr_an_code;

cr_status cr_an_batch(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint64_t *raw, uint32_t stride,
                      const cr_pair_tuning *tuning, double *kernel_ms);

/* A batch plan uploads the work list once and captures zero-fill, the pair
 * kernel and the copy-back in a hipGraph; cr_batch_plan_run replays it.  The
 * plan owns its device memory until destroyed.  Plans of one context must not
 * run concurrently. */
typedef struct cr_batch_plan cr_batch_plan;
cr_status cr_batch_plan_create(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint32_t stride,
                               const cr_pair_tuning *tuning, cr_batch_plan **out);
cr_status cr_batch_plan_run(cr_batch_plan *plan, uint64_t *raw, double *elapsed_ms);
cr_status cr_batch_plan_destroy(cr_batch_plan *plan);

/* Autotuner: times candidate (rows per lane, items per wave) settings of the
 * batched pair kernel on the given codes (each candidate runs the whole batch
 * `repeats` times) and returns the fastest in *best. */
cr_status cr_autotune_batch(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint32_t repeats,
                            cr_pair_tuning *best, double *best_ms);

/* --- super-A search ------------------------------------------------------------------
 * The search drivers over many A, with every exhaustive histogram computed in
 * one batched launch (cr_an_batch) and the drivers' sequential bookkeeping
 * replayed on the host in A order:
 *   style 0: ancoding_search_super_A of distance_distribution/src/main.cpp:131-177
 *            (-S h: odd A in (2^(h-1), 2^h); stdout lines "A: c[minb] = mincb"
 *            for each new best, then the summary line; codes with k > 16 use
 *            the reference grid with m1, m2 as the driver does);
 *   style 1: ancoding_search_super_A(_by_file) of distance_distribution_super_a/src/main.cpp:194-297
 *            (odd A in [A_start, A_end] or the list As; the CSV file body:
 *            "Total Time [s]" line, device line, one process_result row per A
 *            (super_a/src/globals.cpp:39-116) with the running-best columns).
 * *text receives a malloc'ed NUL-terminated string (free with cr_free).
 * Per-A times in the output are the batch time apportioned by pair count. */
typedef struct cr_supera_request {
    uint32_t style;     /* 0 or 1 */
    uint32_t k;         /* data bits */
    uint32_t h_bits;    /* style 0: the -S value */
    int32_t with_grid;  /* style 1: 0 exhaustive, 1/2 reference grid; style 0: grid dimension for k > 16 */
    uint64_t A_start, A_end;  /* style 1 range */
    const uint64_t *As;       /* style 1 candidate list (used when nAs > 0) */
    uint32_t nAs;
    int32_t verbose;
    uint64_t m1, m2;          /* grid sizes */
} cr_supera_request;
cr_status cr_supera_search(cr_context *ctx, const cr_supera_request *req, char **text, double *device_s);
void cr_free(void *p);

/* --- ANcandidates ----------------------------------------------------------------------
 * For every odd A with exactly abits bits (2^(abits-1) < A < 2^abits) the
 * minimum over d in [1, 2^k) of popcount(c XOR 3c), c = A*d (the NAF weight;
 * worker, ANcandidates/ANcandidates.cpp:35-160); *best = the maximum of these
 * minima, cands = the A attaining it in increasing order (the reference's
 * order is nondeterministic; ANcandidates_check.cpp sorts before comparing),
 * *ncands their number (at most cap written).  k + abits <= 62. */
cr_status cr_an_candidates(cr_context *ctx, uint32_t k, uint32_t abits, uint64_t *best, uint64_t *cands,
                           uint64_t cap, uint64_t *ncands, double *kernel_ms);

/* --- serving: continuous batching ---------------------------------------------------
 * A request loop that keeps histogram state on the device.  A scheduler
 * thread runs iterations: it admits queued requests while their device
 * footprint fits the HBM budget and a slot is free, then launches one pass of
 * the pair kernel over a slice of every active AN request (exhaustive and
 * integer-grid) and one pass of the Hamming kernel over a slice of every
 * active Monte Carlo request, so requests that arrive while others run join at
 * the next iteration and their work shares the device.  Results are
 * independent of the mix: integer histograms are order-independent sums and
 * Monte Carlo words come from fixed Philox counters.  No reference
 * counterpart (the reference runs one code per process call). */
typedef enum cr_job_kind {
    CR_JOB_AN_EXHAUSTIVE = 0, /* raw histogram of the code A*d, d < 2^k (as cr_an_pairs) */
    CR_JOB_AN_GRID = 1,       /* integer grid: words A*(i*stride), i*stride < 2^k, unordered pairs (as cr_an_pairs
                                 with multiplier A*stride and ceil(2^k/stride) words) */
    CR_JOB_HAMMING_MC = 2     /* Extended Hamming weights of `samples` Philox words (as cr_hamming_weights, CR_HAM_MC) */
} cr_job_kind;

typedef struct cr_job {
    uint32_t kind;    /* cr_job_kind */
    uint32_t k;       /* data bits */
    uint64_t A;       /* AN jobs */
    uint64_t stride;  /* AN grid */
    uint64_t samples; /* Hamming Monte Carlo */
    uint64_t seed;
    uint32_t stream;
    uint32_t reserved;
} cr_job;

typedef struct cr_server_config {
    uint64_t hbm_budget_bytes; /* sum of device footprints of active requests (0 = 1 GiB) */
    uint32_t max_active;       /* active request slots (0 = 64) */
    uint32_t rows_per_lane;    /* pair kernel R (0 = 8) */
    double iteration_ms;       /* target device time of one iteration (0 = 50 ms) */
} cr_server_config;

typedef struct cr_job_result {
    uint32_t done;            /* 1 when finished */
    uint32_t status;          /* cr_status of the request */
    uint32_t nbins;           /* bins written */
    uint32_t iterations;      /* iterations the request took part in */
    double submit_s, admit_s, finish_s; /* seconds on the server's clock */
    uint64_t units;           /* pairs (AN) or samples (Hamming) */
} cr_job_result;

typedef struct cr_server_stats {
    uint64_t submitted, completed, cancelled, rejected;
    uint64_t iterations;
    uint64_t active, queued;
    uint64_t footprint_bytes;  /* device bytes held by active requests */
    double busy_device_ms;     /* summed device time of the iterations */
} cr_server_stats;

typedef struct cr_server cr_server;
cr_status cr_server_create(cr_context *ctx, const cr_server_config *cfg, cr_server **out);
/* Queues a request; *id identifies it.  Fails with CR_ERR_ADMISSION if the
 * request alone exceeds the HBM budget, CR_ERR_INVALID_ARGUMENT if malformed. */
cr_status cr_server_submit(cr_server *s, const cr_job *job, uint64_t *id);
/* Copies the result of a finished request (CR_ERR_BUSY while it runs).  hist
 * may be NULL to query the state only. */
cr_status cr_server_result(cr_server *s, uint64_t id, uint64_t *hist, uint32_t nbins, cr_job_result *res);
/* Blocks until the request finished (or timeout_ms elapsed: CR_ERR_BUSY). */
cr_status cr_server_wait(cr_server *s, uint64_t id, double timeout_ms);
cr_status cr_server_cancel(cr_server *s, uint64_t id);
cr_status cr_server_stats_get(cr_server *s, cr_server_stats *out);
/* Seconds on the server's clock (the clock of cr_job_result times). */
double cr_server_now(cr_server *s);
/* Stops the scheduler (cancelling unfinished requests) and frees the server. */
cr_status cr_server_destroy(cr_server *s);

#ifdef __cplusplus
}
#endif

#endif /* CODING_RELIABILITY_H_ */
