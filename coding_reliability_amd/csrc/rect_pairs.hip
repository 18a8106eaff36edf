// Rectangle pair histogram: rows x explicit column list (gfx950).
//
// hist[b] += #{(i, s) : popcount(row_i ^ col_s) == b} over all rows i < nrows
// and columns s < ncols (ordered pairs, self-pairs included).  Rows are either
// generated (row_i = A*(row0+i), truncated to 32 bits when mask32) or read
// from a table; columns always come from a table and are read with scalar
// loads (the column index is wave-uniform), so a pair costs xor + bcnt + LDS
// address + ds_add_u32 as in the exhaustive kernel.
//
// This is the execution scheme for the reference's grid approximation
// (dancoding_grid_1d_shared / dancoding_grid_2d,
// distance_distribution/src/an_coding_grid.cu:60-140): 1D = every code word
// against the sample list, 2D = grid words against the sample list.
#include <algorithm>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

// Rows whose high dwords are all equal (h, wave-uniform): the high-dword
// distance popcount(h ^ v_hi) is one scalar per column, folded into the
// popcount of the low dwords as its accumulate operand (3 VALU per pair, as
// for 32-bit words).  Generated rows A*i of one row block share their high
// dword except where A*i crosses a multiple of 2^32.
template <int R>
__device__ __forceinline__ void rect_cols_hu(uint32_t *lane_col, const uint32_t (&wl)[R], uint32_t h,
                                             const uint64_t *__restrict__ c, uint32_t n) {
    uint32_t q = 0;
    for (; q + 4 <= n; q += 4) {
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = uniform64(c[q + u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t vl = (uint32_t)v[u];
            const uint32_t hd = uniform32((uint32_t)__popc(h ^ (uint32_t)(v[u] >> 32)));
#pragma unroll
            for (int r = 0; r < R; ++r) lds_inc(lane_col + (__popc(wl[r] ^ vl) + hd) * kStride);
        }
    }
    for (; q < n; ++q) {
        const uint64_t v = uniform64(c[q]);
        const uint32_t hd = uniform32((uint32_t)__popc(h ^ (uint32_t)(v >> 32)));
#pragma unroll
        for (int r = 0; r < R; ++r) lds_inc(lane_col + (__popc(wl[r] ^ (uint32_t)v) + hd) * kStride);
    }
}

template <int R, bool W2>
__device__ __forceinline__ void rect_cols(uint32_t *lane_col, const uint64_t (&w)[R], const uint64_t *__restrict__ cols,
                                          uint64_t j0, uint32_t n) {
    uint32_t wl[R], wh[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        wl[r] = (uint32_t)w[r];
        wh[r] = (uint32_t)(w[r] >> 32);
    }
    const uint64_t *c = cols + j0;
    if (W2) {
        const uint32_t h = uniform32(wh[0]);
        bool same = true;
#pragma unroll
        for (int r = 0; r < R; ++r) same = same && wh[r] == h;
        if (__all(same)) {
            rect_cols_hu<R>(lane_col, wl, h, c, n);
            return;
        }
    }
    uint32_t q = 0;
    for (; q + 4 <= n; q += 4) {
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = uniform64(c[q + u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t vl = (uint32_t)v[u], vh = (uint32_t)(v[u] >> 32);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                uint32_t d = __popc(wl[r] ^ vl);
                if (W2) d += __popc(wh[r] ^ vh);
                lds_inc(lane_col + d * kStride);
            }
        }
    }
    for (; q < n; ++q) {
        const uint64_t v = uniform64(c[q]);
        const uint32_t vl = (uint32_t)v, vh = (uint32_t)(v >> 32);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint32_t d = __popc(wl[r] ^ vl);
            if (W2) d += __popc(wh[r] ^ vh);
            lds_inc(lane_col + d * kStride);
        }
    }
}

// Partial last row block: rows beyond nrows go to the dummy row.
template <int R>
__device__ void rect_cols_masked(uint32_t *lane_col, const uint64_t (&w)[R], const bool (&ok)[R],
                                 const uint64_t *__restrict__ cols, uint64_t j0, uint32_t n, uint32_t dummy) {
    for (uint32_t q = 0; q < n; ++q) {
        const uint64_t v = uniform64(cols[j0 + q]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint32_t d = (uint32_t)__popcll(w[r] ^ v);
            lds_inc(lane_col + (ok[r] ? d : dummy) * kStride);
        }
    }
}

template <int R, bool W2>
__global__ __launch_bounds__(64) void rect_kernel(uint64_t A, uint64_t row0, uint64_t nrows,
                                                  const uint64_t *__restrict__ rows,
                                                  const uint64_t *__restrict__ cols, uint64_t ncols, uint32_t CW,
                                                  uint64_t CC, uint64_t G, uint64_t RBlocks, uint64_t total_items,
                                                  unsigned long long *__restrict__ counter,
                                                  unsigned long long *__restrict__ out, uint32_t nbins,
                                                  uint32_t first_bin, uint32_t lds_rows, uint32_t mask32) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, lds_rows);
    uint32_t *const lane_col = lds + lane;
    const uint32_t dummy = lds_rows - 1;
    constexpr uint64_t RB = 64ull * R;
    const uint64_t wmask = mask32 ? 0xFFFFFFFFull : ~0ull;
    uint64_t pending = 0;
    bool dirty = false;

    for (uint64_t t = first_item(); t < total_items; t = following_item(counter, t)) {
        // item = (group of G row blocks, column chunk)
        const uint64_t rbg = uniform64(t / CC), cc = uniform64(t - (t / CC) * CC);
        const uint64_t j0 = cc * CW;
        const uint64_t j1 = j0 + CW < ncols ? j0 + CW : ncols;
        const uint32_t n = uniform32((uint32_t)(j1 - j0));
        const uint64_t rb_end = rbg * G + G < RBlocks ? rbg * G + G : RBlocks;
        for (uint64_t rb = rbg * G; rb < rb_end; ++rb) {
            if ((pending + n) * R >= 0xFFFFFFFFull) {
                flush_rows(lds, first_bin, nbins, out);
                pending = 0;
            }
            pending += n;
            dirty = true;

            uint64_t w[R];
            bool ok[R];
            bool full = true;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t i = rb * RB + (uint64_t)r * 64 + lane;
                ok[r] = i < nrows;
                full = full && ((rb * RB + (uint64_t)r * 64 + 63) < nrows);
                const uint64_t ii = ok[r] ? i : 0;
                w[r] = (rows ? rows[ii] : A * (row0 + ii)) & wmask;
            }
            if (full)
                rect_cols<R, W2>(lane_col, w, cols, j0, n);
            else
                rect_cols_masked<R>(lane_col, w, ok, cols, j0, n, dummy);
        }
    }
    if (dirty) flush_rows(lds, first_bin, nbins, out);
}

template <int R, bool W2>
const void *rect_ptr() {
    return reinterpret_cast<const void *>(&rect_kernel<R, W2>);
}

}  // namespace

cr_status launch_rect(const RectLaunch &L, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1, int cus) {
    constexpr int R = 8;
    const void *k = L.w2 ? rect_ptr<R, true>() : rect_ptr<R, false>();
    if (L.nrows == 0 || L.ncols == 0) return CR_OK;
    if (L.lds_rows < 2 || L.lds_rows > 66) return fail(CR_ERR_INVALID_ARGUMENT, "bad histogram rows %u", L.lds_rows);
    const uint64_t RB = 64ull * R;
    const uint64_t RBlocks = (L.nrows + RB - 1) / RB;
    // column chunk: about 16 items per resident wave, at most 2^22 columns
    const size_t lds = (size_t)L.lds_rows * dev::kStride * sizeof(uint32_t);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, lds) != hipSuccess || per_cu <= 0)
        return fail(CR_ERR_DEVICE, "occupancy query failed for the rect kernel");
    const uint64_t waves_cap = (uint64_t)per_cu * cus;
    const uint64_t target = 16 * waves_cap;
    uint64_t CC = (target + RBlocks - 1) / RBlocks;
    if (CC < 1) CC = 1;
    uint64_t CW = (L.ncols + CC - 1) / CC;
    if (CW > (1u << 22)) CW = 1u << 22;
    if (CW < 1) CW = 1;
    CC = (L.ncols + CW - 1) / CW;
    // many more row blocks than waves (every word against a short sample
    // list): G row blocks per item keep the items near the target count, so
    // the work-queue counter (one atomic per item) stays off the critical path
    const uint64_t G = RBlocks > target ? RBlocks / target : 1;
    const uint64_t items = (RBlocks + G - 1) / G * CC;
    uint32_t waves = (uint32_t)std::min<uint64_t>(waves_cap, items);
    CR_TRY(fill_zero_u64(L.counter, 1, stream));
    if (ev0) CR_HIP(hipEventRecord(ev0, stream));
    uint64_t a_A = L.A, a_row0 = L.row0, a_nrows = L.nrows, a_ncols = L.ncols, a_CC = CC, a_items = items, a_G = G,
             a_RB = RBlocks;
    const uint64_t *a_rows = L.rows, *a_cols = L.cols;
    uint32_t a_CW = (uint32_t)CW, a_nbins = L.nbins, a_first = L.first_bin, a_lr = L.lds_rows, a_m32 = L.mask32;
    unsigned long long *a_counter = L.counter, *a_out = L.out;
    void *args[] = {&a_A, &a_row0, &a_nrows, (void *)&a_rows, (void *)&a_cols, &a_ncols, &a_CW, &a_CC, &a_G,
                    &a_RB, &a_items, &a_counter, &a_out, &a_nbins, &a_first, &a_lr, &a_m32};
    CR_HIP(hipLaunchKernel(k, dim3(waves), dim3(64), args, lds, stream));
    if (ev1) CR_HIP(hipEventRecord(ev1, stream));
    return CR_OK;
}

}  // namespace cr
