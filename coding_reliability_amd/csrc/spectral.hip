// Exhaustive AN distance distribution by the Walsh-Hadamard spectrum of the
// code (gfx950).
//
// Same result as the pair kernel (raw[d] = #{i<j : popcount(A*i ^ A*j) = d},
// the quantity dancoding_shared accumulates, distance_distribution/src/
// an_coding.cu:50-105), computed without visiting the 2^(2k-1) pairs.  For the
// code C = {A*i : i < 2^k} in {0,1}^n with indicator f and Walsh-Hadamard
// transform F(S) = sum_x f(x) (-1)^(S.x):
//   D(d) = #{(x, y) in C^2 : |x ^ y| = d} = 2^-n sum_w K_d(w) W(w),
//   W(w) = sum_{|S| = w} F(S)^2,
// with K_d the Krawtchouk polynomials of length n (the autocorrelation of f
// has transform F^2; summing it over the weight-d shell gives K_d).
//
// For odd A the low k bits x_lo = A*i mod 2^k run over every residue once
// (i = x_lo * A^-1 mod 2^k), so with x = (h(x_lo), x_lo), S = (s, S_lo):
//   F(s, S_lo) = WHT_k[ (-1)^(s . h(x_lo)) ](S_lo),
// one length-2^k transform of a +-1 sequence per s < 2^(n-k).  The transform
// runs in passes through a device buffer: for k >= 17, index bits 0-15 in a
// 1024-thread workgroup (int32 out), then bits 16-23 in place when k >= 25,
// then the remaining top bits (8 of them on 256 x 64 tiles, 1..7 with a
// thread per column); k = 32 runs its last two stages in 64 bits (values
// reach 2^32); for 14 <= k <= 16, bits 0-7 (int16 out, values within +-2^8),
// then the rest.  The last pass squares and sums
// F^2 by weight |s| + |S_lo|
// exactly (lo/hi 32-bit halves in 64-bit per-workgroup slots).  The host
// applies the Krawtchouk transform in 128-bit integers.
#include <atomic>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kThreads = 256;  // 4 waves; thread t: column t & 63, quarter p = t >> 6
constexpr uint32_t kRows = 256;     // transform length of a pass (8 bits)
constexpr uint32_t kCols = 64;      // contiguous columns (or transforms) per tile
constexpr uint32_t kExch = 4 * 16 * kCols;  // exchange buffer (ints): one round of 16 rows per wave

// In-register butterflies over the 6 low bits of the register index.
__device__ __forceinline__ void wht64(int32_t (&v)[64]) {
#pragma unroll
    for (int h = 1; h < 64; h <<= 1)
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if ((j & h) == 0) {
                const int32_t a = v[j], b = v[j + h];
                v[j] = a + b;
                v[j + h] = a - b;
            }
}

// Row bits 6 and 7 (the quarter p) of a 256-row transform whose rows
// 64 p + j (j < 64) sit in thread quarter p's registers after wht64.  Four
// rounds through E (16 rows per wave each, conflict free: the lane is the
// column) leave rows 64 m + 16 r + 4 p + l in w[16 r + 4 l + m], and the
// butterflies over m finish the transform.
__device__ __forceinline__ void exchange_gather(const int32_t (&w)[64], int32_t (&o)[64], int32_t *E, uint32_t c,
                                                uint32_t p) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) E[(p * 16 + jj) * kCols + c] = w[16 * r + jj];
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int m = 0; m < 4; ++m) o[16 * r + 4 * l + m] = E[(m * 16 + 4 * p + l) * kCols + c];
        __syncthreads();
    }
}
__device__ __forceinline__ void exchange_m(int32_t (&w)[64], int32_t *E, uint32_t c, uint32_t p) {
    int32_t o[64];
    exchange_gather(w, o, E, c, p);
#pragma unroll
    for (int g = 0; g < 16; ++g) {
        const int32_t *q = o + 4 * g;
        const int32_t a0 = q[0] + q[1], a1 = q[0] - q[1], a2 = q[2] + q[3], a3 = q[2] - q[3];
        w[4 * g] = a0 + a2;
        w[4 * g + 2] = a0 - a2;
        w[4 * g + 1] = a1 + a3;
        w[4 * g + 3] = a1 - a3;
    }
}

// Row held in w[i] after exchange_m, for thread quarter p.
__device__ __forceinline__ uint32_t held_row(int i, uint32_t p) {
    return 64u * (i & 3) + 16u * (i >> 4) + 4u * p + ((i >> 2) & 3);
}

// High part of the codeword with low bits x_lo + 1, from h = high part at
// x_lo (both below A): the word is A * i with i = x_lo * Ainv mod 2^k, and
// A * Ainv = 1 + tc * 2^k, so stepping x_lo adds 1 + tc * 2^k to the word
// and subtracts A * 2^k when i wraps; the low k bits x_lo + 1 do not carry
// (a thread's 64 consecutive x_lo start 64-aligned).  Hence h' = h + tc - A
// exactly when h + tc >= A, i.e. h' = (h + tc) mod A (tc < A since Ainv <
// 2^k): min(t, t - A) in 32 bits (h < 2^(n-k) <= 2^24 and tc < A < 2^25
// for the supported shapes, n <= min(2k, 48), so t never wraps).
__device__ __forceinline__ uint32_t next_high(uint32_t h, uint32_t tc, uint32_t A) {
    const uint32_t t = h + tc;
    return min(t, t - A);
}

// Pass 1: generates the +-1 sequence of chunk s = s0 + blockIdx.y and
// transforms bits 0..7 of the index.  Tile: 64 transforms (index bits 8..)
// of 256 contiguous elements, thread column c = transform.  Element x_lo of
// the chunk holds (-1)^parity(s & h(x_lo)), h(x_lo) = (A * (x_lo * Ainv mod
// 2^k)) >> k < 2^(n-k) <= 2^24, advanced along x_lo by next_high.  The
// transformed tile is staged in LDS (transform c, row r at
// c * 256 + (r ^ c): conflict free both ways) and written row by row.
template <class Out>
__global__ __launch_bounds__(kThreads) void spec_gen(Out *__restrict__ X, uint32_t k, uint32_t A, uint32_t Ainv,
                                                     uint32_t tc, uint32_t s0) {
    __shared__ int32_t E[kCols * kRows];  // 64 KB; the exchange uses its first kExch entries
    const uint32_t t = threadIdx.x, c = t & 63, p = t >> 6;
    const uint32_t s = s0 + blockIdx.y;
    const uint32_t mask = (1u << k) - 1;
    const uint32_t g = blockIdx.x * kCols + c;  // transform: index bits 8..
    const uint32_t x0 = (g << 8) | (p << 6);    // first of this thread's 64 consecutive x_lo
    const uint32_t i = (uint32_t)((uint64_t)x0 * Ainv) & mask;
    uint32_t h = (uint32_t)(((uint64_t)A * i) >> k);
    int32_t v[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        v[j] = 1 - 2 * (int32_t)(__popc(h & s) & 1);
        h = next_high(h, tc, A);
    }
    wht64(v);
    exchange_m(v, E, c, p);
#pragma unroll
    for (int q = 0; q < 64; ++q) E[c * kRows + (held_row(q, p) ^ c)] = v[q];
    __syncthreads();
    Out *out = X + ((uint64_t)blockIdx.y << k) + ((uint64_t)blockIdx.x * kCols << 8);
#pragma unroll 8
    for (uint32_t r = 0; r < kCols; ++r) out[r * kRows + t] = (Out)E[r * kRows + (t ^ r)];
}

// First pass over index bits 0..15 for k = 24 and 32: workgroup (G, chunk) of 1024
// threads transforms the 2^16 elements G * 2^16 + e.  Thread (wave w, lane l)
// generates e = w * 4096 + l * 64 + j (j < 64) and transforms bits 0-5 in
// registers; bits 6-11 (the lane) after a transpose within the wave through
// LDS as int8 (values within +-2^6; four per dword, dword (l, m) of wave w at
// w * 1024 + l * 16 + (m ^ (l >> 2)): conflict free both ways); bits 12-15
// (the wave) after an exchange through LDS as int16 (values within +-2^12) in
// two rounds of 32 rows, thread (w, l) ending with rows 4w .. 4w+3 of every
// wave; the result (within +-2^16) is written as int32, 256 bytes per row.
constexpr uint32_t kThreads16 = 1024;
__global__ __launch_bounds__(kThreads16) void spec_gen16(int32_t *__restrict__ X, uint32_t k, uint32_t A,
                                                         uint32_t Ainv, uint32_t tc, uint32_t s0) {
    const uint32_t mask = k >= 32 ? 0xFFFFFFFFu : (1u << k) - 1;
    __shared__ int32_t L[16384];  // 64 KB
    const uint32_t t = threadIdx.x, l = t & 63, w = t >> 6;
    const uint32_t s = s0 + blockIdx.y;
    const uint32_t G = blockIdx.x;
    const uint32_t x0 = (G << 16) | (w << 12) | (l << 6);
    const uint32_t i = (uint32_t)((uint64_t)x0 * Ainv) & mask;
    uint32_t h = (uint32_t)(((uint64_t)A * i) >> k);
    int32_t v[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        v[j] = 1 - 2 * (int32_t)(__popc(h & s) & 1);
        h = next_high(h, tc, A);
    }
    wht64(v);  // bits 0-5; |v| <= 64
    // transpose within the wave: element (row l, column j) -> thread (w, j)
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const uint32_t d = (uint32_t)(v[4 * m] & 0xFF) | ((uint32_t)(v[4 * m + 1] & 0xFF) << 8) |
                           ((uint32_t)(v[4 * m + 2] & 0xFF) << 16) | ((uint32_t)(v[4 * m + 3] & 0xFF) << 24);
        L[w * 1024 + l * 16 + (m ^ (l >> 2))] = (int32_t)d;
    }
    __syncthreads();
    {
        const int8_t *L8 = reinterpret_cast<const int8_t *>(L);
        const uint32_t m = l >> 2, b = l & 3;  // this thread's column j = l: dword m, byte b
#pragma unroll
        for (int r = 0; r < 64; ++r) v[r] = L8[4 * (w * 1024 + r * 16 + (m ^ ((uint32_t)r >> 2))) + b];
    }
    wht64(v);  // bits 6-11 (v[r]: row r); |v| <= 2^12
    __syncthreads();
    // exchange across waves: thread (w, l) gathers rows 4w + rr (rr < 4) of all 16 waves
    int16_t *L16 = reinterpret_cast<int16_t *>(L);  // [wave][32 rows][64 lanes] per round
    int32_t u[64];                                   // u[4 * src + rr]
#pragma unroll
    for (int hr = 0; hr < 2; ++hr) {
#pragma unroll
        for (int rr = 0; rr < 32; ++rr) L16[(w * 32 + rr) * 64 + l] = (int16_t)v[32 * hr + rr];
        __syncthreads();
        if ((w >> 3) == (uint32_t)hr) {
#pragma unroll
            for (int src = 0; src < 16; ++src)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) u[4 * src + rr] = L16[(src * 32 + 4 * (w & 7) + rr) * 64 + l];
        }
        __syncthreads();
    }
    // bits 12-15: butterflies over src (bits 2-5 of the u index)
#pragma unroll
    for (int hb = 4; hb < 64; hb <<= 1)
#pragma unroll
        for (int q = 0; q < 64; ++q)
            if ((q & hb) == 0) {
                const int32_t a = u[q], c2 = u[q + hb];
                u[q] = a + c2;
                u[q + hb] = a - c2;
            }
    int32_t *out = X + ((uint64_t)blockIdx.y << k) + ((uint64_t)G << 16);  // |u| <= 2^16
#pragma unroll
    for (int src = 0; src < 16; ++src)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) out[src * 4096 + (4 * w + rr) * 64 + l] = u[4 * src + rr];
}

// Element index of (row, column) in tile tau of a pass over bits
// [tbit, tbit + 8): idx = tau_hi << (tbit + 8) | row << tbit | tau_lo << 6 | col.
__device__ __forceinline__ uint64_t pass_index(uint32_t tau, uint32_t tbit, uint32_t row, uint32_t col) {
    const uint32_t lo_bits = tbit - 6;
    const uint64_t tau_lo = tau & ((1u << lo_bits) - 1), tau_hi = tau >> lo_bits;
    return (tau_hi << (tbit + 8)) | ((uint64_t)row << tbit) | (tau_lo << 6) | col;
}

// Loads the 256 x 64 tile (rows strided by 2^tbit, 64 contiguous columns)
// and transforms its rows; w then holds rows held_row(i, p).
// Rows p * 64 + j (j < 64) of tile tau, column c, into w[j].  For int32 data
// one buffer resource per wave (p is the wave index) covers the wave's 64
// rows, under 2^32 bytes for tbit <= 24: the row offset is a scalar and the
// column one 32-bit VGPR, so no 64-bit address per load occupies VGPRs.
template <class In>
__device__ __forceinline__ void load_rows(const In *chunk, uint32_t tau, uint32_t tbit, uint32_t c, uint32_t p,
                                          int32_t (&w)[64]) {
    if constexpr (sizeof(In) == 4) {
        const uint32_t pw = __builtin_amdgcn_readfirstlane(p);
        const In *row0 = chunk + pass_index(tau, tbit, pw * 64, 0);
        const uint32_t row_bytes = 4u << tbit;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<In *>(row0), 0, (int)(63u * row_bytes + 64u * 4u), 0x00020000);
#pragma unroll
        for (int j = 0; j < 64; ++j)
            w[j] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(c * 4u), (int)(j * row_bytes), 0);
    } else {
#pragma unroll
        for (int j = 0; j < 64; ++j) w[j] = (int32_t)chunk[pass_index(tau, tbit, p * 64 + j, c)];
    }
}

template <class In>
__device__ __forceinline__ void pass_rows(const In *chunk, uint32_t tau, uint32_t tbit, int32_t *E,
                                          int32_t (&w)[64]) {
    const uint32_t t = threadIdx.x, c = t & 63, p = t >> 6;
    load_rows<In>(chunk, tau, tbit, c, p, w);
    wht64(w);
    exchange_m(w, E, c, p);
}

// Middle pass for k = 32: index bits [tbit, tbit + 8) of chunk blockIdx.y,
// int32 in place (values within +-2^24; every read of a tile precedes the
// workgroup barriers that precede its writes).
__global__ __launch_bounds__(kThreads) void spec_mid(int32_t *__restrict__ X, uint32_t k, uint32_t tbit) {
    __shared__ int32_t E[kExch];
    const uint32_t t = threadIdx.x, c = t & 63, p = t >> 6;
    int32_t *chunk = X + ((uint64_t)blockIdx.y << k);
    int32_t w[64];
    pass_rows<int32_t>(chunk, blockIdx.x, tbit, E, w);
#pragma unroll
    for (int q = 0; q < 64; ++q) chunk[pass_index(blockIdx.x, tbit, held_row(q, p), c)] = w[q];
}

// Adds the workgroup's LDS bins (lo, hi halves per weight) to its slot: four
// uint64 per weight, (lo sum, carries of it, hi sum, carries of it), so a
// slot holds 128-bit sums however many launches add to it.  One workgroup per
// slot and launch: plain read-modify-write.
__device__ __forceinline__ void flush_slot(const unsigned long long *bins, unsigned long long *slot,
                                           uint32_t nbins) {
    for (uint32_t b = threadIdx.x; b < 2 * nbins; b += blockDim.x) {
        const unsigned long long x = bins[b], old = slot[2 * b];
        slot[2 * b] = old + x;
        if (old + x < old) slot[2 * b + 1] += 1;
    }
}

// F^2 into the 32-bit halves lo, hi (|f| <= 2^32: F^2 = 2^64 is hi = 2^32).
// Branch free: |f| = t * 2^32 + a with t in {0, 1} and a = 0 when t = 1, so
// F^2 = a^2 + t * 2^64, one 32 x 32 -> 64-bit product.
__device__ __forceinline__ void add_square(int64_t f, unsigned long long &lo, unsigned long long &hi) {
    const unsigned long long m = (unsigned long long)(f < 0 ? -f : f);
    const uint32_t a = (uint32_t)m, t = (uint32_t)(m >> 32);
    const unsigned long long q = (unsigned long long)a * a;
    lo += q & 0xFFFFFFFFull;
    hi += ((unsigned long long)t << 32) | (q >> 32);
}

// Last pass: bits [tbit, tbit + 8) of every chunk of the batch (tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... of nchunks * tiles_per_chunk), then
// W[|s| + |S_lo|] += F^2.  The weight of held_row(i, p) is |p| + |i >> 4| +
// |(i >> 2) & 3| + |i & 3| over the tile's base |s| + |tau| + |col|, so each
// thread sums into 7 pairs of registers per tile (low and high 32-bit halves
// of F^2: a half-sum per tile is below 2^38); the workgroup adds them to LDS
// bins, and at the end to its own slot (4 * nbins uint64 at slots +
// blockIdx.x * 4 * nbins, see flush_slot).  WIDE
// (k = 32): the values reach +-2^32, so the last two butterfly stages run in
// 64 bits.
template <class In, bool WIDE>
__global__ __launch_bounds__(kThreads) void spec_last(const In *__restrict__ X, uint32_t k, uint32_t tbit,
                                                      uint64_t s0, uint32_t nchunks, uint32_t tiles_per_chunk,
                                                      unsigned long long *__restrict__ slots, uint32_t nbins) {
    __shared__ int32_t E[kExch];
    __shared__ unsigned long long bins[2 * 64];
    const uint32_t t = threadIdx.x, c = t & 63, p = t >> 6;
    for (uint32_t b = t; b < 2 * 64; b += kThreads) bins[b] = 0;
    __syncthreads();
    const uint64_t ntiles = (uint64_t)nchunks * tiles_per_chunk;
    const uint32_t lo_bits = tbit - 6;
    for (uint64_t q = blockIdx.x; q < ntiles; q += gridDim.x) {
        const uint32_t ch = (uint32_t)(q / tiles_per_chunk), tau = (uint32_t)(q - (uint64_t)ch * tiles_per_chunk);
        const In *src = X + ((uint64_t)ch << k);
        unsigned long long lo[7] = {0, 0, 0, 0, 0, 0, 0}, hi[7] = {0, 0, 0, 0, 0, 0, 0};
        int32_t w[64];
        if (WIDE) {
            load_rows<In>(src, tau, tbit, c, p, w);  // k = 32: rows 2^26 bytes apart
            wht64(w);                                // |w| <= 2^30
            int32_t o[64];
            exchange_gather(w, o, E, c, p);
#pragma unroll
            for (int g = 0; g < 16; ++g) {
                const int64_t q0 = o[4 * g], q1 = o[4 * g + 1], q2 = o[4 * g + 2], q3 = o[4 * g + 3];
                const int64_t a0 = q0 + q1, a1 = q0 - q1, a2 = q2 + q3, a3 = q2 - q3;
                const int64_t f[4] = {a0 + a2, a1 + a3, a0 - a2, a1 - a3};  // rows m = 0..3
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int o_ = __builtin_popcount(g >> 2) + __builtin_popcount(g & 3) + __builtin_popcount(m);
                    add_square(f[m], lo[o_], hi[o_]);
                }
            }
        } else {
            // |F| <= 2^24: a thread's 64 squares stay below 2^54 in one register pair
            pass_rows<In>(src, tau, tbit, E, w);
            unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 64; ++i) {
                const int64_t f = w[i];
                acc[__builtin_popcount(i >> 4) + __builtin_popcount((i >> 2) & 3) + __builtin_popcount(i & 3)] +=
                    (unsigned long long)(f * f);
            }
#pragma unroll
            for (int o_ = 0; o_ < 7; ++o_) {
                lo[o_] = acc[o_] & 0xFFFFFFFFull;
                hi[o_] = acc[o_] >> 32;
            }
        }
        const uint32_t base = (uint32_t)__popcll(s0 + ch) + __popc(tau & ((1u << lo_bits) - 1)) +
                              __popc(tau >> lo_bits) + __popc(c) + __popc(p);
#pragma unroll
        for (int o_ = 0; o_ < 7; ++o_) {
            atomicAdd(&bins[2 * (base + o_)], lo[o_]);
            atomicAdd(&bins[2 * (base + o_) + 1], hi[o_]);
        }
    }
    __syncthreads();
    flush_slot(bins, slots + (uint64_t)blockIdx.x * 4 * nbins, nbins);
}

// ---------------------------------------------------------------------------
// k = 17 .. 24 through int16 (2 bytes of traffic per element each way): the
// first pass transforms index bits 0 .. B1-1, B1 = k - 10 (values within
// +-2^B1 fit int16), of two chunks at once, the last pass the remaining 10
// bits.
//
// The passes transform the parities q(x) = parity(s & h(x)) in {0, 1} rather
// than the signs (-1)^q = 1 - 2q: the transform of 1 - 2q is 2^k delta(S) -
// 2 Q(S), Q the transform of q, so F^2 / 4 = Q^2 except at S = 0, where it is
// (Q(0) - 2^(k-1))^2.  The last pass sums F^2 / 4 and the host multiplies the
// weight sums by 4.  After B1 <= 14 stages |Q| <= 2^14, within int16.
//
// spec_gen_pairs<W, B1>: workgroup (G, y) of W waves (W = 2^(B1-12) for k =
// 22 .. 24, W = 1 below) takes the 2^SPAN elements G * 2^SPAN + e (SPAN =
// max(B1, 12)) of chunks sa = s0 + 2y (even) and sa + 1,
// each dword of its working set holding the two chunks' values as int16
// halves (packed v_pk_add_u16 / v_pk_sub_i16 butterflies: one instruction per
// two values).  Since sa + 1 differs from sa in bit 0 only, parity((sa + 1) &
// h) = parity(sa & h) ^ (h & 1).  Thread (wave w, lane l) generates e = w *
// 4096 + l * 64 + j (j < 64): bits 0-5 in registers; bits 6 .. min(B1, 12)-1
// after a transpose within the wave (a lane swap, then 32 x 32 transposes
// through LDS rows of 33 dwords); bits 12 .. B1-1 after an exchange across the
// waves in two rounds of 32 registers.  For B1 < 12 the untransformed bits B1
// .. 11 are the low bits of the row x_hi.  Each chunk's result is written as int16 in the
// layout Y[c][S_lo >> 6][x_hi][S_lo & 63] (x_hi < 1024), so that a tile of
// the last pass -- all x_hi of 32 consecutive S_lo -- is one contiguous run of
// 64-byte rows.
typedef short v2s __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));  // a buffer descriptor in four SGPRs
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2s, a) + __builtin_bit_cast(v2s, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2s, a) - __builtin_bit_cast(v2s, b));
}
// Packed butterflies over the register-index bits of strides [H0, H1).
template <int H0, int H1>
__device__ __forceinline__ void wht_pk(uint32_t (&v)[64]) {
#pragma unroll
    for (int h = H0; h < H1; h <<= 1)
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if ((j & h) == 0) {
                const uint32_t a = v[j], b = v[j + h];
                v[j] = pk_add(a, b);
                v[j + h] = pk_sub(a, b);
            }
}

template <int W, int B1>
__global__ __launch_bounds__(64 * W, 2) void spec_gen_pairs(int16_t *__restrict__ Y, uint32_t k, uint32_t A,
                                                            uint32_t Ainv, uint32_t tc, uint32_t s0) {
    // B1 index bits transformed here; a workgroup covers 2^SPAN elements, so
    // for W = 1 and B1 < 12 its top 12 - B1 bits are left to the last pass
    constexpr uint32_t SPAN = W == 4 ? 14 : (W == 2 ? 13 : 12);
    static_assert(W == 1 ? (B1 >= 7 && B1 <= 12) : B1 == (int)SPAN, "spec_gen_pairs shape");
    constexpr int QB = 32 / W;  // registers per wave and exchange round
    // per wave a 64 x 33-dword transpose buffer, then (W > 1) the exchange of
    // W * 32 * 64 dwords: 33.8 KB at W = 4, four workgroups per CU
    __shared__ uint32_t L[W * 64 * 33];
    const uint32_t t = threadIdx.x, l = t & 63, w = t >> 6;
    const uint32_t mask = (1u << k) - 1;
    const uint32_t G = blockIdx.x;
    const uint32_t sa = s0 + 2 * blockIdx.y;
    const uint32_t x0 = (G << SPAN) | (w << 12) | (l << 6);
    const uint32_t i = (uint32_t)((uint64_t)x0 * Ainv) & mask;
    uint32_t h = (uint32_t)(((uint64_t)A * i) >> k);
    uint32_t v[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        const uint32_t pa = __popc(h & sa);
        v[j] = (pa | ((pa ^ h) << 16)) & 0x00010001u;  // parities of chunks sa (low) and sa+1 (high)
        h = next_high(h, tc, A);
    }
    wht_pk<1, 64>(v);  // bits 0-5; |v| <= 2^6
    // transpose within the wave (afterwards lane l' holds as register r what
    // lane r held as register l').  First v_permlane32_swap exchanges lane
    // bit 5 with register bit 5: for j < 32 the upper 32 lanes of v[j] trade
    // places with the lower 32 lanes of v[j + 32], so lane (a, m) (a = bit 5)
    // then holds as register (b, r) what lane (b, m) held as register (a, r).
    // The rest is a 32 x 32 transpose of the low five bits per register half
    // b, through rows of 33 dwords: lane (a, m) stores register (b, r) at
    // R[33 (32 a + r) + m] and reads it back from R[33 (32 a + m) + r]; the
    // register index is an immediate offset and each half-wave's banks, (r +
    // m) mod 32 up to a constant, are distinct.  A wave's LDS accesses execute
    // in order, so no barrier separates them.  The buffer is half the size of
    // a direct 64 x 65 transpose, which doubles the resident workgroups.
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const auto sw = __builtin_amdgcn_permlane32_swap(v[j], v[j + 32], false, false);
        v[j] = sw[0];
        v[j + 32] = sw[1];
    }
    uint32_t *R = L + w * (64 * 33);
    const uint32_t a = l >> 5, m = l & 31;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int r = 0; r < 32; ++r) R[(32 * a + r) * 33 + m] = v[32 * b + r];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < 32; ++r) v[32 * b + r] = R[(32 * a + m) * 33 + r];
        __builtin_amdgcn_wave_barrier();
    }
    // bits 6 .. min(B1, 12) - 1 (v[r]: bits 6-11 = r, bits 0-5 = l); |v| <= 2^min(B1, 12)
    wht_pk<1, (B1 >= 12 ? 64 : (1 << (B1 - 6)))>(v);
    if (W > 1) {
        // bits 12 .. B1-1 (the wave index): after round hr, v[32 hr + QB q +
        // rr] holds register 32 hr + QB w + rr of wave q
#pragma unroll
        for (int hr = 0; hr < 2; ++hr) {
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < 32; ++rr) L[(w * 32 + rr) * 64 + l] = v[32 * hr + rr];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < W; ++q)
#pragma unroll
                for (int rr = 0; rr < QB; ++rr) v[32 * hr + QB * q + rr] = L[(q * 32 + QB * w + rr) * 64 + l];
        }
        wht_pk<QB, 32>(v);  // over q; |v| <= 2^B1
    }
    // stores through one buffer resource per chunk (2^k int16): the lane part
    // of the offset in a VGPR, the rest wave-uniform
    const uint32_t X = 1u << (k - B1);
    int16_t *Ya = Y + ((uint64_t)(2 * blockIdx.y) << k);
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(Ya, 0, (int)(2u << k), 0x00020000);
    // chunk sa + 1 takes the high halves: buffer_store_short_d16_hi writes
    // bits 16-31 of the register as they are (no shift), through a
    // descriptor of the same form as make_buffer_rsrc's
    const uint64_t yb = (uint64_t)(Ya + ((uint64_t)1 << k));
    v4i rb;
    rb.x = (int)(uint32_t)yb;
    rb.y = (int)(uint32_t)(yb >> 32);
    rb.z = (int)(2u << k);
    rb.w = 0x00020000;
    const uint32_t wu = __builtin_amdgcn_readfirstlane(w);
    // the workgroup's first row x_hi; for W = 1 and B1 < 12 register r holds
    // S_lo >> 6 = r mod 2^(B1-6) of row x_hi + (r >> (B1-6))
    const int voff = (int)(((G << (SPAN - B1)) * 64 + l) * 2);
#pragma unroll
    for (int hr = 0; hr < 2; ++hr)
#pragma unroll
        for (int q = 0; q < W; ++q)
#pragma unroll
            for (int rr = 0; rr < QB; ++rr) {
                const uint32_t r = 32u * hr + QB * wu + rr;
                const uint32_t sb = W == 1 ? (r & ((1u << (B1 - 6 < 6 ? B1 - 6 : 6)) - 1)) : (((uint32_t)q << 6) | r);
                const uint32_t xr = W == 1 && B1 < 12 ? r >> (B1 - 6) : 0;
                const int soff = (int)((sb * X + xr) * 128u);  // bytes of sb * X + xr rows of 64
                const uint32_t x = v[32 * hr + QB * q + rr];
                __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(x & 0xFFFFu), ra, voff, soff, 0);
                asm volatile("buffer_store_short_d16_hi %0, %1, %2, %3 offen"
                             :
                             : "v"(x), "v"(voff), "s"(rb), "s"(soff)
                             : "memory");
            }
}

// spec_last10: tile (c, sb, half) = rows x_hi < 1024 and the 32 columns S_lo =
// sb * 64 + half * 32 + cc of chunk c (64 KB of int16).  Thread (row group g
// < 32, column pair c2 < 16) loads rows g * 32 + j (j < 32) of columns 2 c2 and
// 2 c2 + 1 as one dword per row (buffer loads: tile base in the resource, row
// offset a scalar), transforms x_hi bits 0-4 of both columns in registers,
// then bits 5-9 after one exchange per column through LDS (every thread
// stores its 32 registers, then gathers register g of the 32 row groups of its
// column pair), and sums F^2 / 4 by weight |s| + |S_lo| + |S_hi| as spec_last
// sums F^2.  The exchange rows are 528 dwords apart, so the two row groups of
// each half-wave fall on opposite halves of the banks in both directions.
// Tiles are visited so that the two column halves of a row segment go to
// workgroups b and b + 8, which run on the same XCD (workgroups are dealt to
// the 8 XCDs in turn) at the same time.
constexpr uint32_t kLast10Threads = 512;
constexpr uint32_t kLast10Pitch = 528;  // dwords between the exchange rows of two registers
__device__ __forceinline__ void wht32(int32_t (&v)[32]) {
#pragma unroll
    for (int h = 1; h < 32; h <<= 1)
#pragma unroll
        for (int j = 0; j < 32; ++j)
            if ((j & h) == 0) {
                const int32_t a = v[j], b = v[j + h];
                v[j] = a + b;
                v[j + h] = a - b;
            }
}
__global__ __launch_bounds__(kLast10Threads) void spec_last10(const int16_t *__restrict__ Y, uint32_t k, uint64_t s0,
                                                              uint32_t nchunks, unsigned long long *__restrict__ slots,
                                                              uint32_t nbins) {
    __shared__ uint32_t E[31 * kLast10Pitch + 32 * 16];  // 67.5 KB
    // weight bins in 8 copies: bins[(w * 2 + half) * 8 + (c2 & 7)] (lo, hi
    // halves of F^2); the lanes of an add share an address at most 8-fold
    __shared__ unsigned long long bins[64 * 2 * 8];  // 8 KB
    const uint32_t t = threadIdx.x, c2 = t & 15, g = t >> 4;
    for (uint32_t b = t; b < 64 * 2 * 8; b += kLast10Threads) bins[b] = 0;
    __syncthreads();
    const uint32_t X = 1024;                  // rows x_hi: the first pass took index bits 0 .. k-11
    const uint32_t sbits = k - 16;            // bits of sb
    const uint64_t ntiles = (uint64_t)nchunks << (k - 15);
    const int voff = (int)((g * 32 * 64 + 2 * c2) * 2);
    for (uint64_t u = blockIdx.x; u < ntiles; u += gridDim.x) {
        const uint64_t jx = u >> 3;
        const uint32_t half = (uint32_t)(jx & 1);
        const uint64_t P = ((jx >> 1) << 3) | (u & 7);
        const uint32_t c = (uint32_t)(P >> sbits), sb = (uint32_t)(P & ((1u << sbits) - 1));
        const int16_t *tile = Y + ((uint64_t)c << k) + (uint64_t)sb * X * 64 + half * 32;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int16_t *>(tile), 0, (int)(((X - 1) * 64 + 32) * 2), 0x00020000);
        int32_t va[32], vb[32];  // columns 2 c2 and 2 c2 + 1
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const uint32_t d = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, j * 128, 0);
            va[j] = (int32_t)(int16_t)(d & 0xFFFFu);
            vb[j] = (int32_t)d >> 16;
        }
        wht32(va);  // x_hi bits 0-4; |v| <= 2^19
        wht32(vb);
        // v[gg] afterwards: x_hi bits 5-9 = gg, transformed bits 0-4 = g
#pragma unroll
        for (int col = 0; col < 2; ++col) {
            int32_t(&v)[32] = col ? vb : va;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 32; ++j) E[j * kLast10Pitch + g * 16 + c2] = (uint32_t)v[j];
            __syncthreads();
#pragma unroll
            for (int gg = 0; gg < 32; ++gg) v[gg] = (int32_t)E[g * kLast10Pitch + gg * 16 + c2];
        }
        wht32(va);  // x_hi bits 5-9
        wht32(vb);
        // v = Q: F / 2 = -Q, except at S = 0 (thread 0, register 0 of column
        // 0 of the tiles with sb = 0 and half = 0), where F / 2 = 2^(k-1) - Q
        if (sb == 0 && half == 0 && t == 0) va[0] -= 1 << (k - 1);
        // |F / 2| <= 2^23: a thread's 64 squares stay below 2^53 in one
        // register pair; column 2 c2 + 1 has one more set bit than 2 c2
        unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int64_t fa = va[q], fb = vb[q];
            acc[__builtin_popcount(q)] += (unsigned long long)(fa * fa);
            acc[__builtin_popcount(q) + 1] += (unsigned long long)(fb * fb);
        }
        const uint32_t base = (uint32_t)__popcll(s0 + c) + __popc(sb) + half + __popc(c2) + __popc(g);
#pragma unroll
        for (int o_ = 0; o_ < 7; ++o_) {
            atomicAdd(&bins[((base + o_) * 2) * 8 + (c2 & 7)], acc[o_] & 0xFFFFFFFFull);
            atomicAdd(&bins[((base + o_) * 2 + 1) * 8 + (c2 & 7)], acc[o_] >> 32);
        }
    }
    __syncthreads();
    // the 8 copies into the first (one thread per (w, half))
    for (uint32_t b = t; b < 2 * nbins; b += kLast10Threads) {
        unsigned long long x = 0;
        for (uint32_t q = 0; q < 8; ++q) x += bins[b * 8 + ((q + b) & 7)];
        bins[b * 8] = x;
    }
    __syncthreads();
    unsigned long long *slot = slots + (uint64_t)blockIdx.x * 4 * nbins;
    for (uint32_t b = t; b < 2 * nbins; b += kLast10Threads) {
        const unsigned long long x = bins[b * 8], old = slot[2 * b];
        slot[2 * b] = old + x;
        if (old + x < old) slot[2 * b + 1] += 1;
    }
}

// Narrow last pass: the R < 8 top index bits [tbit, tbit + R) of every chunk,
// tiles of 2^R rows (strided by 2^tbit) x 256 contiguous columns, thread t
// owning column t and all 2^R rows in registers (no exchange).  Element
// (row, column) of tile tau is tau_hi << (tbit + R) | row << tbit |
// tau_lo << 8 | column, tau_lo being tbit - 8 bits.  F^2 is summed as in
// spec_last, by the popcount of the row over the base |s| + |tau| + |col|;
// k <= 31 here, so F^2 < 2^62.
template <class In, int R>
__global__ __launch_bounds__(kThreads) void spec_last_narrow(const In *__restrict__ X, uint32_t k, uint32_t tbit,
                                                             uint64_t s0, uint32_t nchunks,
                                                             uint32_t tiles_per_chunk,
                                                             unsigned long long *__restrict__ slots,
                                                             uint32_t nbins) {
    constexpr int NR = 1 << R;
    __shared__ unsigned long long bins[2 * 64];
    const uint32_t t = threadIdx.x;
    for (uint32_t b = t; b < 2 * 64; b += kThreads) bins[b] = 0;
    __syncthreads();
    const uint64_t ntiles = (uint64_t)nchunks * tiles_per_chunk;
    const uint32_t lo_bits = tbit - 8;
    for (uint64_t q = blockIdx.x; q < ntiles; q += gridDim.x) {
        const uint32_t ch = (uint32_t)(q / tiles_per_chunk), tau = (uint32_t)(q - (uint64_t)ch * tiles_per_chunk);
        const uint64_t tau_lo = tau & ((1u << lo_bits) - 1), tau_hi = tau >> lo_bits;
        const In *src = X + ((uint64_t)ch << k) + (tau_hi << (tbit + R)) + (tau_lo << 8) + t;
        int64_t v[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) v[r] = (int64_t)src[(uint64_t)r << tbit];
#pragma unroll
        for (int h = 1; h < NR; h <<= 1)
#pragma unroll
            for (int j = 0; j < NR; ++j)
                if ((j & h) == 0) {
                    const int64_t a = v[j], b = v[j + h];
                    v[j] = a + b;
                    v[j + h] = a - b;
                }
        unsigned long long lo[R + 1], hi[R + 1];
#pragma unroll
        for (int o_ = 0; o_ <= R; ++o_) lo[o_] = hi[o_] = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const unsigned long long sq = (unsigned long long)(v[r] * v[r]);
            lo[__builtin_popcount(r)] += sq & 0xFFFFFFFFull;
            hi[__builtin_popcount(r)] += sq >> 32;
        }
        const uint32_t base = (uint32_t)__popcll(s0 + ch) + __popc(tau) + __popc(t);
#pragma unroll
        for (int o_ = 0; o_ <= R; ++o_) {
            atomicAdd(&bins[2 * (base + o_)], lo[o_]);
            atomicAdd(&bins[2 * (base + o_) + 1], hi[o_]);
        }
    }
    __syncthreads();
    flush_slot(bins, slots + (uint64_t)blockIdx.x * 4 * nbins, nbins);
}

template <class In>
cr_status launch_narrow(uint32_t R, const In *X, uint32_t k, uint32_t tbit, uint64_t s0, uint32_t nchunks,
                        unsigned long long *slots, uint32_t nslots, uint32_t nbins, hipStream_t stream) {
    const uint32_t tiles = 1u << (k - R - 8);
#define CR_NARROW(RR)                                                                                            \
    case RR:                                                                                                     \
        hipLaunchKernelGGL((spec_last_narrow<In, RR>), dim3(nslots), dim3(kThreads), 0, stream, X, k, tbit, \
                           s0, nchunks, tiles, slots, nbins);                                                    \
        break;
    switch (R) {
        CR_NARROW(1) CR_NARROW(2) CR_NARROW(3) CR_NARROW(4) CR_NARROW(5) CR_NARROW(6) CR_NARROW(7)
    default:
        return fail(CR_ERR_UNSUPPORTED, "spectral path: narrow pass of %u bits", R);
    }
#undef CR_NARROW
    CR_HIP(hipGetLastError());
    return CR_OK;
}

template <class In>
const void *narrow_kernel(uint32_t R) {
    switch (R) {
    case 1: return reinterpret_cast<const void *>(&spec_last_narrow<In, 1>);
    case 2: return reinterpret_cast<const void *>(&spec_last_narrow<In, 2>);
    case 3: return reinterpret_cast<const void *>(&spec_last_narrow<In, 3>);
    case 4: return reinterpret_cast<const void *>(&spec_last_narrow<In, 4>);
    case 5: return reinterpret_cast<const void *>(&spec_last_narrow<In, 5>);
    case 6: return reinterpret_cast<const void *>(&spec_last_narrow<In, 6>);
    case 7: return reinterpret_cast<const void *>(&spec_last_narrow<In, 7>);
    default: return nullptr;
    }
}

}  // namespace

size_t spectral_tile_count(uint32_t k) { return (size_t)1 << (k - 14); }

uint32_t spectral_last_pass_per_cu(uint32_t layout, uint32_t k) {
    // the kernel that launch_spectral_batch runs last, and a cache index per
    // kernel (the query runs once per kernel and process: every gfx950 device
    // gives the same answer)
    if (k < 14 || k > 32) return 4;
    const void *f = nullptr;
    int threads = (int)kThreads, slot = 0;
    if (layout == kSpectralInt16Pairs) {
        f = reinterpret_cast<const void *>(&spec_last10);
        threads = (int)kLast10Threads;
    } else if (k < 16) {
        f = narrow_kernel<int16_t>(k - 8);
        slot = 1 + (int)(k - 8);  // 7 .. 8
    } else if (k == 16) {
        f = reinterpret_cast<const void *>(&spec_last<int16_t, false>);
        slot = 9;
    } else if (k < 24 || (k > 24 && k < 32)) {
        f = narrow_kernel<int32_t>(k < 24 ? k - 16 : k - 24);
        slot = 10 + (int)(k < 24 ? k - 16 : k - 24);  // 11 .. 17
    } else if (k == 24) {
        f = reinterpret_cast<const void *>(&spec_last<int32_t, false>);
        slot = 18;
    } else {
        f = reinterpret_cast<const void *>(&spec_last<int32_t, true>);
        slot = 19;
    }
    static std::atomic<int> cache[20];
    const int known = cache[slot].load(std::memory_order_relaxed);
    if (known > 0) return (uint32_t)known;
    int n = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, threads, 0) != hipSuccess || n <= 0) return 4;
    cache[slot].store(n, std::memory_order_relaxed);
    return (uint32_t)n;
}

cr_status launch_spectral_batch(void *X, uint32_t layout, uint32_t k, uint64_t A, uint32_t Ainv, uint64_t tc,
                                uint64_t s0, uint32_t nchunks, unsigned long long *slots, uint32_t nslots,
                                uint32_t nbins, hipStream_t stream) {
    if (k < 14 || k > 32) return fail(CR_ERR_UNSUPPORTED, "spectral path: k=%u outside [14, 32]", k);
    if (nbins > 64) return fail(CR_ERR_UNSUPPORTED, "spectral path: %u weights", nbins);
    if ((A >> 32) || (tc >> 32) || (s0 + nchunks) > (1ull << 32))
        return fail(CR_ERR_UNSUPPORTED, "spectral path: multiplier or transform count beyond 32 bits");
    if (layout == kSpectralInt16Pairs) {
        // X: nchunks * 2^k int16 after index bits 0 .. k-11, in pairs of transforms
        // spec_last10 pairs the tiles in groups of 16 (8 row blocks x 2
        // column halves): the tile count must be a multiple of 16
        if (k < 17 || k > 24 || (nchunks & 1) || (s0 & 1) || (nslots & 15) ||
            (((uint64_t)nchunks << (k - 15)) & 15))
            return fail(CR_ERR_INVALID_ARGUMENT, "spectral int16 pairs: k=%u, %u transforms from %llu, %u slots", k,
                        nchunks, (unsigned long long)s0, nslots);
        int16_t *y = static_cast<int16_t *>(X);
        // k = 24 / 23: four / two waves over index bits 0-13 / 0-12; k <= 22:
        // one wave over bits 0 .. k-11 of 2^12 elements (1024 rows each)
        const dim3 grid(k >= 23 ? 1024 : 1u << (k - 12), nchunks / 2);
        const uint32_t a32 = (uint32_t)A, tc32 = (uint32_t)tc, s32 = (uint32_t)s0;
#define CR_GEN(WW, BB)                                                                                  \
    hipLaunchKernelGGL((spec_gen_pairs<WW, BB>), grid, dim3(64 * WW), 0, stream, y, k, a32, Ainv, tc32, s32); \
    break;
        switch (k) {
        case 24: CR_GEN(4, 14)
        case 23: CR_GEN(2, 13)
        case 22: CR_GEN(1, 12)
        case 21: CR_GEN(1, 11)
        case 20: CR_GEN(1, 10)
        case 19: CR_GEN(1, 9)
        case 18: CR_GEN(1, 8)
        default: CR_GEN(1, 7)
        }
#undef CR_GEN
        CR_HIP(hipGetLastError());
        hipLaunchKernelGGL(spec_last10, dim3(nslots), dim3(kLast10Threads), 0, stream, (const int16_t *)y, k, s0,
                           nchunks, slots, nbins);
        CR_HIP(hipGetLastError());
        return CR_OK;
    }
    const uint32_t tiles = (uint32_t)spectral_tile_count(k);
    if (k <= 16) {
        // X: nchunks * 2^k int16 after index bits 0-7 (tiles of 64 transforms: k >= 14)
        int16_t *x16 = static_cast<int16_t *>(X);
        hipLaunchKernelGGL(spec_gen<int16_t>, dim3(tiles, nchunks), dim3(kThreads), 0, stream, x16, k, (uint32_t)A,
                           Ainv, (uint32_t)tc, (uint32_t)s0);
        CR_HIP(hipGetLastError());
        if (k < 16)
            return launch_narrow<int16_t>(k - 8, x16, k, 8, s0, nchunks, slots, nslots, nbins, stream);
        hipLaunchKernelGGL((spec_last<int16_t, false>), dim3(nslots), dim3(kThreads), 0, stream,
                           (const int16_t *)x16, k, 8u, s0, nchunks, tiles, slots, nbins);
        CR_HIP(hipGetLastError());
        return CR_OK;
    }
    // X: nchunks * 2^k int32 after index bits 0-15
    int32_t *x32 = static_cast<int32_t *>(X);
    hipLaunchKernelGGL(spec_gen16, dim3(1u << (k - 16), nchunks), dim3(kThreads16), 0, stream, x32, k, (uint32_t)A,
                       Ainv, (uint32_t)tc, (uint32_t)s0);
    CR_HIP(hipGetLastError());
    if (k < 24) return launch_narrow<int32_t>(k - 16, x32, k, 16, s0, nchunks, slots, nslots, nbins, stream);
    if (k == 24) {
        hipLaunchKernelGGL((spec_last<int32_t, false>), dim3(nslots), dim3(kThreads), 0, stream,
                           (const int32_t *)x32, k, 16u, s0, nchunks, tiles, slots, nbins);
        CR_HIP(hipGetLastError());
        return CR_OK;
    }
    hipLaunchKernelGGL(spec_mid, dim3(tiles, nchunks), dim3(kThreads), 0, stream, x32, k, 16u);
    CR_HIP(hipGetLastError());
    if (k < 32) return launch_narrow<int32_t>(k - 24, x32, k, 24, s0, nchunks, slots, nslots, nbins, stream);
    hipLaunchKernelGGL((spec_last<int32_t, true>), dim3(nslots), dim3(kThreads), 0, stream, (const int32_t *)x32, k,
                       24u, s0, nchunks, tiles, slots, nbins);
    CR_HIP(hipGetLastError());
    return CR_OK;
}

}  // namespace cr
