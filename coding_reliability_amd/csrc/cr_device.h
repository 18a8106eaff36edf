// Device helpers shared by the histogram kernels (gfx950, wave64).
//
// All histogram kernels use the same LDS scheme: one single-wave workgroup
// owns rows of 64 uint32 counters, one row per distance/weight bin and one
// counter per lane (row stride 64 dwords), so every per-pair ds_add_u32 lands
// in the lane's own bank whatever the bin is.  Rows are flushed to 64-bit
// device histograms with one atomic per bin.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cr {
namespace dev {

constexpr uint32_t kStride = 64;  // dwords between LDS histogram rows (one per lane)

// readfirstlane returns int: go through uint32_t so the low dword is not sign-extended
__device__ __forceinline__ uint32_t uniform32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    return ((uint64_t)uniform32((uint32_t)(x >> 32)) << 32) | (uint64_t)uniform32((uint32_t)x);
}

__device__ __forceinline__ void lds_inc(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_clear(uint32_t *lds, uint32_t rows) {
    for (uint32_t b = 0; b < rows; ++b) lds[b * kStride + threadIdx.x] = 0;
    __syncthreads();
}

// Adds every lane column of rows [first, nbins) to out[] and clears the rows.
// Lane l reduces row l (and l+64); at step s it reads column (s+l) mod 64, so
// the 64 lanes of one read hit 64 different banks.
__device__ inline void flush_rows(uint32_t *lds, uint32_t first, uint32_t nbins, unsigned long long *out) {
    __syncthreads();
    const uint32_t lane = threadIdx.x;
    for (uint32_t b = lane; b < nbins; b += 64) {
        uint32_t *row = lds + b * kStride;
        unsigned long long s = 0;
#pragma unroll 8
        for (uint32_t step = 0; step < 64; ++step) {
            const uint32_t l = (step + lane) & 63u;
            s += row[l];
            row[l] = 0;
        }
        if (s && b >= first) atomicAdd(out + b, s);
    }
    __syncthreads();
}

// Histogram copies: a histogram of nbins entries may be kept as `copies`
// copies (a power of two; 0 means 1), hist_copy_stride(nbins) entries apart,
// each on whole 128-byte lines; wave b flushes into copy b mod copies and the
// host sums the copies.  Many waves flushing one histogram otherwise queue
// their atomics on its few cache lines.
constexpr uint32_t kHistCopies = 32;
__host__ __device__ inline uint64_t hist_copy_stride(uint32_t nbins) { return ((uint64_t)nbins + 15) & ~15ull; }
__device__ __forceinline__ unsigned long long *hist_copy(unsigned long long *out, uint32_t nbins, uint32_t copies) {
    const uint32_t c = copies > 1 ? (blockIdx.x & (copies - 1)) : 0;
    return out + (uint64_t)c * hist_copy_stride(nbins);
}

// Work-queue head: lane 0 takes the next item, every lane sees it.
__device__ __forceinline__ uint64_t next_item(unsigned long long *counter) {
    uint64_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(counter, 1ull);
    return uniform64(t);
}

// Item order of the persistent kernels (every launch has at most as many
// waves W = gridDim.x as items): wave b's first item is item b, taken without
// an atomic; its later items come from the work queue at `counter`, whose
// k-th step hands out item W + k, or, with counter == nullptr (a static
// assignment), are b + W, b + 2W, ...  Each queue step is a device-scope
// atomic on one address, serialized at about 12.5 ns on MI355X.
__device__ __forceinline__ uint64_t first_item() { return blockIdx.x; }
__device__ __forceinline__ uint64_t following_item(unsigned long long *counter, uint64_t t) {
    return counter ? (uint64_t)gridDim.x + next_item(counter) : t + gridDim.x;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of the 4x32
// Philox S-box, multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.  c is the counter on input, the output on return.
// A round is two 32x32->64 products (v_mad_u64_u32) and two three-input XORs
// (v_bitop3_b32 with truth table 0x96, the key in a scalar register).
__device__ __forceinline__ void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t x0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c[1], k0, 0x96);
        const uint32_t x2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c[3], k1, 0x96);
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = x0;
        c[2] = x2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Tag in the fourth counter word of the library's Monte Carlo streams.
constexpr uint32_t kMcTag = 0x43520001u;

}  // namespace dev
}  // namespace cr
