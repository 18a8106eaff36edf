// Exhaustive pair-distance histogram kernel for AN codes (gfx950).
//
// Computes raw[b] = #{i<j : popcount(A*i ^ A*j) == b}, the quantity the
// reference kernel dancoding_shared accumulates
// (distance_distribution/src/an_coding.cu:50-105), with a different execution
// scheme:
//   * one wave per workgroup, persistent, pulling items from a global queue;
//   * R row codewords per lane live in VGPRs (A*i computed, never loaded);
//   * the column codeword is wave-uniform and advances in SGPRs (v += A), so a
//     pair costs one v_xor + one v_bcnt (+ bcnt accumulate of the constant
//     high-dword distance for 33..64-bit words) + one LDS address mad + one
//     ds_add_u32 into the lane's private column of the histogram;
//   * LDS histogram rows are 64 dwords (one per lane), so the bank of every
//     per-pair atomic is fixed by the lane whatever the distance is: the hot
//     loop never bank-conflicts; the flush reads each row in a lane-rotated
//     order, which is conflict free too; counters are flushed to the 64-bit
//     device histogram with one atomic per bin per flush, before any 32-bit
//     lane counter can wrap.
#include <atomic>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {

namespace {

using namespace dev;
constexpr uint32_t kUnroll = 4;  // columns per unrolled step

template <int R>
struct Rows {
    uint64_t w[R];  // codeword A*i_r
    uint64_t i[R];  // row index i_r
};

// Diagonal block: columns [j0, j1) are below (b+1)*RB, pairs with j <= i go to
// the dummy row.  Generic 64-bit path; at most RB columns per row block.
template <int R>
__device__ void pairs_masked(uint32_t *lane_col, const Rows<R> &rw, uint64_t A, uint64_t j0, uint64_t j1,
                             uint32_t dummy) {
    uint64_t v = A * j0;
    for (uint64_t j = j0; j < j1; ++j, v += A) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint32_t c = (uint32_t)__popcll(rw.w[r] ^ v);
            c = j > rw.i[r] ? c : dummy;
            lds_inc(lane_col + c * kStride);
        }
    }
}

// All codewords < 2^32.
template <int R>
__device__ void pairs_w1(uint32_t *lane_col, const Rows<R> &rw, uint64_t A, uint64_t j0, uint64_t j1) {
    uint32_t w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = (uint32_t)rw.w[r];
    const uint32_t a = (uint32_t)A;
    uint32_t v = uniform32((uint32_t)(A * j0));
    uint32_t n = uniform32((uint32_t)(j1 - j0));  // items span < 2^32 columns
    for (; n >= kUnroll; n -= kUnroll) {
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds_inc(lane_col + __popc(w[r] ^ v) * kStride);
            v += a;
        }
    }
    for (; n; --n) {
#pragma unroll
        for (int r = 0; r < R; ++r) lds_inc(lane_col + __popc(w[r] ^ v) * kStride);
        v += a;
    }
}

// 64-bit codewords, A < 2^32: within a segment of columns the high dword of
// A*j is constant, so its distance to each row's high dword is computed once
// per segment and fed to v_bcnt as the accumulate operand.
template <int R>
__device__ void pairs_w2seg(uint32_t *lane_col, const Rows<R> &rw, uint64_t A, uint64_t j0, uint64_t j1) {
    uint32_t wlo[R], whi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        wlo[r] = (uint32_t)rw.w[r];
        whi[r] = (uint32_t)(rw.w[r] >> 32);
    }
    const uint32_t a = (uint32_t)A;
    uint64_t j = j0;
    while (j < j1) {
        const uint64_t v64 = uniform64(A * j);
        const uint32_t vhi = (uint32_t)(v64 >> 32);
        uint64_t je = j1;
        if (vhi != 0xFFFFFFFFu) {
            // first column whose high dword exceeds vhi
            const uint64_t next = ((((uint64_t)vhi + 1) << 32) + A - 1) / A;
            if (next < je) je = next;
        }
        je = uniform64(je);
        uint32_t cr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) cr[r] = __popc(whi[r] ^ vhi);
        uint32_t v = (uint32_t)v64;
        uint32_t n = uniform32((uint32_t)(je - j));
        for (; n >= kUnroll; n -= kUnroll) {
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
#pragma unroll
                for (int r = 0; r < R; ++r) lds_inc(lane_col + (__popc(wlo[r] ^ v) + cr[r]) * kStride);
                v += a;
            }
        }
        for (; n; --n) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds_inc(lane_col + (__popc(wlo[r] ^ v) + cr[r]) * kStride);
            v += a;
        }
        j = je;
    }
}

// 64-bit codewords, generic: two xors and two bcnts per pair.
template <int R>
__device__ void pairs_w2gen(uint32_t *lane_col, const Rows<R> &rw, uint64_t A, uint64_t j0, uint64_t j1) {
    uint64_t v = uniform64(A * j0);
    uint32_t n = uniform32((uint32_t)(j1 - j0));
    for (; n >= kUnroll; n -= kUnroll) {
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds_inc(lane_col + __popcll(rw.w[r] ^ v) * kStride);
            v += A;
        }
    }
    for (; n; --n) {
#pragma unroll
        for (int r = 0; r < R; ++r) lds_inc(lane_col + __popcll(rw.w[r] ^ v) * kStride);
        v += A;
    }
}

__global__ void fill_zero_kernel(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <int R>
__global__ __launch_bounds__(64) void pair_kernel(const PairJob *__restrict__ jobs, uint32_t njobs,
                                                  uint64_t total_items, unsigned long long *__restrict__ counter,
                                                  uint32_t lds_rows) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, lds_rows);
    uint32_t *const lane_col = lds + lane;
    const uint32_t dummy = lds_rows - 1;
    constexpr uint64_t RB = 64ull * R;

    uint32_t cur = 0xFFFFFFFFu;  // job whose counts sit in LDS
    uint64_t pending = 0;        // columns accumulated since the last flush

    for (uint64_t t = first_item(); t < total_items; t = following_item(counter, t)) {
        // job lookup: current job first, else binary search over item_begin
        uint32_t jb;
        if (cur != 0xFFFFFFFFu && t >= jobs[cur].item_begin && t < jobs[cur].item_begin + jobs[cur].items) {
            jb = cur;
        } else {
            uint32_t lo = 0, hi = njobs - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (jobs[mid].item_begin <= t) lo = mid; else hi = mid - 1;
            }
            jb = uniform32(lo);
        }
        const PairJob &job = jobs[jb];
        const uint64_t A = job.A, N = job.N, NC = job.NC, Q = job.Q;
        const uint64_t tl = t - job.item_begin + job.item_offset;
        // row block: largest b with prefix(b) <= tl
        uint64_t lo = 0, hi = job.T - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if (pair_prefix(mid, NC, Q) <= tl) lo = mid; else hi = mid - 1;
        }
        const uint64_t b = uniform64(lo);
        const uint64_t c = (Q == 1 ? b : b / Q) + (tl - pair_prefix(b, NC, Q));
        const uint64_t CW = Q * RB;
        const uint64_t row0 = b * RB, row_end = row0 + RB;
        uint64_t j0 = c * CW;
        if (j0 < row0) j0 = row0;
        uint64_t j1 = (c + 1) * CW;
        if (j1 > N) j1 = N;
        j0 = uniform64(j0);
        j1 = uniform64(j1);

        if (jb != cur || (pending + (j1 - j0)) * R >= 0xFFFFFFFFull) {
            if (cur != 0xFFFFFFFFu)
                flush_rows(lds, 0, jobs[cur].nbins, hist_copy(jobs[cur].out, jobs[cur].nbins, jobs[cur].out_copies));
            cur = jb;
            pending = 0;
        }
        pending += j1 - j0;

        Rows<R> rw;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rw.i[r] = row0 + (uint64_t)r * 64 + lane;
            rw.w[r] = A * rw.i[r];
        }
        const uint64_t jd1 = j1 < row_end ? j1 : row_end;
        if (j0 < jd1) pairs_masked<R>(lane_col, rw, A, j0, jd1, dummy);
        const uint64_t jr0 = j0 > row_end ? j0 : row_end;
        if (jr0 < j1) {
            if (job.mode == kModeW1)
                pairs_w1<R>(lane_col, rw, A, jr0, j1);
            else if (job.mode == kModeW2Seg)
                pairs_w2seg<R>(lane_col, rw, A, jr0, j1);
            else
                pairs_w2gen<R>(lane_col, rw, A, jr0, j1);
        }
    }
    if (cur != 0xFFFFFFFFu)
        flush_rows(lds, 0, jobs[cur].nbins, hist_copy(jobs[cur].out, jobs[cur].nbins, jobs[cur].out_copies));
}

template <int R>
const void *pair_kernel_ptr() {
    return reinterpret_cast<const void *>(&pair_kernel<R>);
}

const void *pair_kernel_for(uint32_t R) {
    switch (R) {
    case 2: return pair_kernel_ptr<2>();
    case 4: return pair_kernel_ptr<4>();
    case 8: return pair_kernel_ptr<8>();
    case 16: return pair_kernel_ptr<16>();
    default: return nullptr;
    }
}

}  // namespace

uint32_t pair_max_distance(uint64_t A, uint64_t N) {
    const unsigned __int128 m = (unsigned __int128)(N - 1) * A;
    uint32_t bits = 0;
    for (unsigned __int128 x = m; x; x >>= 1) ++bits;
    return bits;
}

uint32_t pair_mode(uint64_t A, uint64_t N) {
    const uint32_t bits = pair_max_distance(A, N);
    if (bits <= 32) return kModeW1;
    // high dword of A*j stays constant for about 2^32/A consecutive columns
    if (A <= (1ull << 22)) return kModeW2Seg;
    return kModeW2Gen;
}

void pair_job_geometry(PairJob &job, uint32_t R, uint32_t Q) {
    const uint64_t RB = 64ull * R;
    job.T = (job.N + RB - 1) / RB;
    if (Q == 0) Q = 1;
    job.Q = Q;
    const uint64_t CW = (uint64_t)Q * RB;
    job.NC = (job.N + CW - 1) / CW;
    job.items = pair_prefix(job.T, job.NC, job.Q);
}

uint32_t pair_waves_per_cu(uint32_t R, uint32_t lds_rows) {
    // one occupancy query per (R, rows): the answer is a property of the
    // kernel and of gfx950's CU, the only target
    static std::atomic<uint32_t> cache[17][67];
    const void *k = pair_kernel_for(R);
    if (!k || lds_rows > 66) return 0;
    const uint32_t hit = cache[R][lds_rows].load(std::memory_order_relaxed);
    if (hit) return hit;
    int blocks = 0;
    const size_t lds = (size_t)lds_rows * kStride * sizeof(uint32_t);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64, lds) != hipSuccess) return 0;
    cache[R][lds_rows].store((uint32_t)blocks, std::memory_order_relaxed);
    return (uint32_t)blocks;
}

cr_status fill_zero_u64(unsigned long long *p, size_t n, hipStream_t stream) {
    if (n == 0) return CR_OK;
    const uint32_t threads = 256;
    uint32_t blocks = (uint32_t)((n + threads - 1) / threads);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(threads), 0, stream, p, n);
    CR_HIP(hipGetLastError());
    return CR_OK;
}

cr_status launch_pairs(const PairJob *d_jobs, uint32_t njobs, uint64_t total_items, unsigned long long *counter,
                       const PairLaunchConfig &cfg, int cus, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
    const void *k = pair_kernel_for(cfg.R);
    if (!k) return fail(CR_ERR_INVALID_ARGUMENT, "rows_per_lane must be 2, 4, 8 or 16 (got %u)", cfg.R);
    if (cfg.lds_rows < 2 || cfg.lds_rows > 66)
        return fail(CR_ERR_INVALID_ARGUMENT, "histogram rows %u outside [2, 66]", cfg.lds_rows);
    if (njobs == 0) return fail(CR_ERR_INVALID_ARGUMENT, "no jobs");
    uint32_t waves = cfg.waves;
    if (waves == 0) {
        const uint32_t per_cu = pair_waves_per_cu(cfg.R, cfg.lds_rows);
        if (per_cu == 0) return fail(CR_ERR_DEVICE, "occupancy query failed for the pair kernel");
        waves = per_cu * (uint32_t)cus;
    }
    if ((uint64_t)waves > total_items) waves = (uint32_t)(total_items ? total_items : 1);
    const size_t lds = (size_t)cfg.lds_rows * kStride * sizeof(uint32_t);
    if (cfg.allow_static && total_items > 0) {
        // at most one wave per CU less than the resident capacity, each
        // taking the same number of items
        counter = nullptr;
        const uint32_t per_cu = pair_waves_per_cu(cfg.R, cfg.lds_rows);
        const uint64_t cap = (uint64_t)(per_cu > 1 ? per_cu - 1 : 1) * cus;
        if (waves > cap) waves = (uint32_t)cap;
        const uint64_t per = (total_items + waves - 1) / waves;
        waves = (uint32_t)((total_items + per - 1) / per);
    }
    if (cfg.zero_counter && counter) CR_TRY(fill_zero_u64(counter, 1, stream));
    if (ev0) CR_HIP(hipEventRecord(ev0, stream));
    const PairJob *a_jobs = d_jobs;
    uint32_t a_njobs = njobs, a_rows = cfg.lds_rows;
    uint64_t a_items = total_items;
    unsigned long long *a_counter = counter;
    void *args[] = {(void *)&a_jobs, (void *)&a_njobs, (void *)&a_items, (void *)&a_counter, (void *)&a_rows};
    CR_HIP(hipLaunchKernel(k, dim3(waves), dim3(64), args, lds, stream));
    if (ev1) CR_HIP(hipEventRecord(ev1, stream));
    return CR_OK;
}

}  // namespace cr
