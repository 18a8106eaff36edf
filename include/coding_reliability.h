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
    CR_ERR_UNSUPPORTED = 2,      /* valid request outside what this build supports */
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

/* --- exhaustive AN ---------------------------------------------------------
 * raw[b] = #{ 0 <= i < j < num_words : popcount(A*i XOR A*j) == b } for
 * b < nbins, computed exactly on the GPU (uint64 counts).
 * Requires A >= 1, num_words >= 1 and (num_words-1)*A < 2^64; nbins must cover
 * every possible distance (bit length of (num_words-1)*A, plus one).
 * With num_words = 2^k this is the histogram accumulated by dancoding_shared
 * (distance_distribution/src/an_coding.cu:50-105) before run_ancoding doubles it
 * (an_coding.cu:287-292).  With A = A0*s and num_words = m it is the
 * integer-stride grid {A0*(i*s) : i < m}.
 * kernel_ms (may be NULL) receives the device time of the pair kernel. */
cr_status cr_an_pairs(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                      double *kernel_ms);

/* Tuning knobs for the pair kernel; 0 = automatic (autotuned table). */
typedef struct cr_pair_tuning {
    uint32_t rows_per_lane;   /* codewords held in registers per lane: 2, 4, 8 or 16 */
    uint32_t chunk_blocks;    /* column chunk in units of the 64*rows row block */
    uint32_t waves;           /* persistent waves launched (0 = fill the device) */
} cr_pair_tuning;
cr_status cr_an_pairs_tuned(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                            const cr_pair_tuning *tuning, double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* CODING_RELIABILITY_H_ */
