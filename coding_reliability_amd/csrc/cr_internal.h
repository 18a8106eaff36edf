// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "coding_reliability.h"

namespace cr {

// Records a message for cr_last_error() on this thread and returns s.
cr_status fail(cr_status s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define CR_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ::cr::fail(e_ == hipErrorOutOfMemory ? CR_ERR_OUT_OF_MEMORY : CR_ERR_DEVICE,      \
                              "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,       \
                              __LINE__);                                                             \
    } while (0)

#define CR_TRY(expr)                                                                                 \
    do {                                                                                             \
        cr_status s_ = (expr);                                                                       \
        if (s_ != CR_OK) return s_;                                                                  \
    } while (0)

// Device memory that grows on demand and is freed with its owner.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    cr_status reserve(size_t want);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host memory that grows on demand.
struct HostBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    cr_status reserve(size_t want);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// ---------------------------------------------------------------------------
// Pair-histogram jobs (exhaustive AN and integer-stride grids).
//
// A job is the triangle {0 <= i < j < N} over codewords A*i.  Rows are cut in
// blocks of RB = 64*R words (R codewords per lane held in registers); the
// columns j >= b*RB of row block b are cut at multiples of CW = Q*RB into
// items.  Item (b, c) covers columns [max(c*CW, b*RB), min((c+1)*CW, N)); the
// part below (b+1)*RB is the diagonal block and is masked to j > i.
// ---------------------------------------------------------------------------
enum PairMode : uint32_t {
    kModeW1 = 0,     // all codewords < 2^32: xor + bcnt per pair
    kModeW2Seg = 1,  // 64-bit codewords, column high dword constant over long segments
    kModeW2Gen = 2,  // 64-bit codewords, generic two-dword xor + bcnt
};

struct PairJob {
    uint64_t A;
    uint64_t N;
    uint64_t item_begin;  // global index of the job's first item in this launch
    uint64_t items;       // items of this job in this launch
    uint64_t item_offset; // local index of the first item (a launch may run a slice)
    uint64_t T;           // row blocks
    uint64_t NC;          // column chunks
    unsigned long long *out;  // device histogram, nbins entries, accumulated with atomics
    uint32_t Q;           // chunk width in row blocks
    uint32_t mode;
    uint32_t nbins;       // bins flushed to out
    uint32_t out_copies;  // copies of the histogram at out (power of two; 0 = 1), hist_copy_stride apart
};
static_assert(sizeof(PairJob) == 80, "PairJob layout");

// sum_{x=0}^{b-1} floor(x/Q); Q = 1 (every small job) without a division
__host__ __device__ inline uint64_t pair_floor_sum(uint64_t b, uint64_t Q) {
    if (Q == 1) return b * (b - 1) / 2;
    const uint64_t q = b / Q, r = b - q * Q;
    return Q * (q * (q - (q > 0 ? 1 : 0)) / 2) + r * q;
}
// items of row blocks [0, b)
__host__ __device__ inline uint64_t pair_prefix(uint64_t b, uint64_t NC, uint64_t Q) {
    return b * NC - pair_floor_sum(b, Q);
}

// Single-code launches below this many pairs run with a static assignment
// (about 100 us of work at 9.6e12 pairs/s: the queue's serialized atomics,
// one per item, would cost more than a static assignment's imbalance).
constexpr long double kStaticPairs = 1e9;

// Tuning for one pair job when the caller leaves a knob at 0: enough items
// (about 16 per resident wave) to balance the persistent waves, rows per lane
// large enough to amortise the per-item set-up.
void auto_pair_tuning(uint64_t N, uint32_t resident_waves, uint32_t &R, uint32_t &Q);

// Looks up (once per device, with that device current) the address of the
// allocation-free entry point's file-scope scratch array on `device`.
cr_status noalloc_setup_device(int device);

// Fills job geometry for R rows per lane and chunk width Q (in row blocks).
void pair_job_geometry(PairJob &job, uint32_t R, uint32_t Q);
// Bit length of the largest codeword (N-1)*A, i.e. the highest possible distance.
uint32_t pair_max_distance(uint64_t A, uint64_t N);
// Chooses the kernel mode for a job.
uint32_t pair_mode(uint64_t A, uint64_t N);

struct PairLaunchConfig {
    uint32_t R = 8;
    uint32_t waves = 0;         // 0 = resident capacity of the device
    uint32_t lds_rows = 0;      // histogram rows in LDS (max nbins + 1 dummy row)
    bool zero_counter = true;   // false: the caller cleared the queue head already
    bool allow_static = false;  // a short launch: static assignment of items to waves, no queue
};

// Launches zero-fill + persistent pair kernel on `stream` for the device job
// array d_jobs (njobs entries, item_begin prefix-summed), with `counter`
// (device u64) as the work queue head.  Histograms jobs[i].out are NOT cleared
// here.  Records ev0/ev1 around the pair kernel when non-null.
cr_status launch_pairs(const PairJob *d_jobs, uint32_t njobs, uint64_t total_items,
                       unsigned long long *counter, const PairLaunchConfig &cfg, int cus,
                       hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1);

// Resident single-wave workgroups per CU for the pair kernel with R rows per
// lane and `lds_rows` histogram rows.
uint32_t pair_waves_per_cu(uint32_t R, uint32_t lds_rows);

// Rectangle histogram rows x column table (see rect_pairs.hip).
struct RectLaunch {
    uint64_t A = 1, row0 = 0, nrows = 0;
    const uint64_t *rows = nullptr;  // device row table, or null for rows A*(row0+i)
    const uint64_t *cols = nullptr;  // device column table
    uint64_t ncols = 0;
    uint32_t mask32 = 0;             // truncate generated rows to 32 bits
    bool w2 = false;                 // some codeword exceeds 32 bits
    unsigned long long *counter = nullptr;
    unsigned long long *out = nullptr;
    uint32_t nbins = 0, first_bin = 0, lds_rows = 0;
};
cr_status launch_rect(const RectLaunch &L, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1, int cus);

// Hamming weight jobs (see hamming.hip).
enum HamSource : uint32_t { kHamRange = 0, kHamStride = 1, kHamList = 2, kHamMc = 3 };
struct HamJob {
    uint64_t first;            // range/stride offset, or first Monte Carlo sample index
    uint64_t stride;
    uint64_t count;            // samples
    const uint64_t *list;      // device list (kHamList)
    uint64_t seed;
    uint64_t item_begin, items;  // global range of this job in the launch
    uint64_t item_offset;        // local index of the first item (slices)
    unsigned long long *out;   // device weight histogram (nbins entries)
    uint64_t masks[7];         // parity masks truncated to k bits, zero when unused
    uint32_t source, k, stream, nbins;
    uint32_t lane_samples;     // samples per lane per work item (power of two in [16, kHamMaxLaneSamples])
    uint32_t out_copies;       // copies of the histogram at out (power of two; 0 = 1), hist_copy_stride apart
};
static_assert(sizeof(HamJob) == 152, "HamJob layout");
constexpr uint32_t kHamMaxLaneSamples = 4096;
constexpr uint32_t kHamMaxBatch = 4096;  // requests per cr_hamming_weights_batch call

// True when num_words = 2^k and cr_an_pairs_spectral supports (A, k) and is
// estimated at under 80 % of the pair kernel's time (measured rates:
// 2.2e-12 s per transformed element plus about 50 us per call, against
// 9.6e12 pairs/s).
bool spectral_preferred(uint64_t A, uint64_t num_words, uint32_t *k);

// One exhaustive AN histogram by the code's spectrum, split into batches of
// transforms (see spectral.hip): shape, device buffer and slot sizes.
struct SpectralPlan {
    uint64_t Ao = 0, tc = 0, chunks = 0, per_batch = 0, batches = 0, buffer_bytes = 0, slot_entries = 0;
    uint32_t k = 0, n = 0, nw = 0, Ainv = 0, nslots = 0;
    uint32_t layout = 0;          // kSpectralInt32 / kSpectralInt16Pairs (see spectral.hip)
    double sec_per_element = 0;   // measured device seconds per transformed element
};
// Pass structures of the spectral path: index bits 0-15 (or 0-7 for k <= 16)
// then the rest through an int32 (int16) buffer; or, for k = 17 .. 24 and at
// least two transforms, bits 0 .. k-11 of two transforms at once then the
// last 10 bits, through int16.
constexpr uint32_t kSpectralInt32 = 0, kSpectralInt16Pairs = 1;
// False when (A, k) is outside the spectral path (then the pair kernel applies).
bool spectral_plan(uint64_t A, uint32_t k, int cus, SpectralPlan *p);
// Launches batches [first_batch, first_batch + nbatches) on stream, with X of
// p.buffer_bytes and slots of p.slot_entries uint64 (zeroed before the first).
cr_status spectral_run(const SpectralPlan &p, void *X, unsigned long long *slots, uint64_t first_batch,
                       uint64_t nbatches, hipStream_t stream);
// raw[] (nbins entries, pair-kernel form) from the host copy of the slots.
cr_status spectral_finish(const SpectralPlan &p, const uint64_t *slots, uint64_t *raw, uint32_t nbins);

// Parity masks of the (N,k) extended Hamming codes, truncated to k bits;
// returns the number of non-overall parity bits.
uint32_t hamming_masks(uint32_t k, uint64_t masks[7]);
// Work items of a job of `count` samples at lane_samples per lane (64 lanes).
uint64_t hamming_items(uint64_t count, uint32_t lane_samples);
// Waves per CU of the Hamming kernel (the larger of the 32- and 64-bit forms).
uint32_t hamming_waves_per_cu(uint32_t lds_rows);
// counter: the work-queue head (zeroed here).  allow_static: a short launch
// of equal items, run with a static assignment of items to waves instead of
// the queue.
cr_status launch_hamming(const HamJob *d_jobs, uint32_t njobs, uint64_t total_items, bool w64,
                         unsigned long long *counter, bool allow_static, uint32_t lds_rows, int cus,
                         hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1);
// A sample range of an AN Monte Carlo launch and its histogram: the pairs
// (word i, sample j) for every word i and j0 <= j < j1 (j0 < j1).
struct McSegment {
    uint64_t j0, j1;
    uint64_t item_begin;      // first work item of the range (set by launch_an_mc)
    uint64_t chunks;          // sample chunks of the range (set by launch_an_mc)
    unsigned long long *out;  // device histogram, nbins entries (zeroed by the caller)
};
constexpr uint32_t kMcMaxRanges = 64;  // sample ranges per AN Monte Carlo launch
// segs (host, nsegs entries, kept alive until the stream is synchronized) is
// completed and copied to d_segs; counter is the work-queue head (zeroed here).
cr_status launch_an_mc(uint64_t A, uint32_t n, uint64_t seed, McSegment *segs, uint32_t nsegs, McSegment *d_segs,
                       unsigned long long *counter, uint32_t nbins, int cus, hipStream_t stream, hipEvent_t ev0,
                       hipEvent_t ev1);

// Zero-fills n 64-bit words on stream (a kernel, so it may be captured in graphs).
cr_status fill_zero_u64(unsigned long long *p, size_t n, hipStream_t stream);

}  // namespace cr

struct cr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cus = 0;
    int clock_mhz = 0;
    uint64_t hbm = 0;
    char name[64] = {0};
    char arch[32] = {0};
    cr::DeviceBuffer jobs;     // PairJob array
    cr::DeviceBuffer counters; // work-queue heads
    cr::DeviceBuffer hist;     // histograms
    cr::DeviceBuffer table_a;  // row / sample tables
    cr::DeviceBuffer table_b;
    cr::HostBuffer host;       // pinned staging
};
