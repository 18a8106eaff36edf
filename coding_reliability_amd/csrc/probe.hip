// Issue-rate probes: the device ceilings the histogram kernels are bound by,
// measured on the device they run on (gfx950, wave64).  Each probe runs one
// instruction mix and nothing else, on every resident single-wave
// workgroup, for a fixed number of iterations; the unit is a VALU
// lane-instruction (the loop bodies were checked in the gfx950 assembly) or
// an LDS lane increment:
//   0 kProbeValu    xor + popcount-accumulate (v_xor_b32, v_bcnt_u32_b32) on
//                   8 independent chains: the pair, grid and exhaustive
//                   Hamming inner loops; 16 per iteration
//   1 kProbePhilox  Philox4x32-10 blocks on 4 independent streams per lane
//                   (20 v_mad_u64_u32 + 20 v_bitop3_b32 per block)
//   2 kProbeLds     wave-wide ds_add_u32 to lane-private LDS counters in the
//                   histogram layout of every kernel (row stride 64 dwords),
//                   the lanes of an instruction in different rows
//   3 kProbeMixed   Philox blocks whose words go through the AN Monte Carlo
//                   pair work (multiply, xor, popcount): 52 per block
//   4 kProbeValu16  as 0 on 16 chains
//   5 kProbeLdsRow  as 2, all lanes of an instruction in one row
//   6 kProbeWrite   streaming 16-byte stores over a 256 MiB buffer (bytes/s)
//   7 kProbeRead    streaming 16-byte loads over the same buffer (bytes/s)
//   8 kProbeLdsTab  ds_read_u8 from 128-byte tables (one dword per LDS bank,
//                   the Hamming syndrome tables' layout), 8 tables read at one
//                   address and combined by three-input XORs: LDS lane-reads/s
#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kProbeValu = 0, kProbePhilox = 1, kProbeLds = 2, kProbeMixed = 3, kProbeValu16 = 4,
                   kProbeLdsRow = 5, kProbeWrite = 6, kProbeRead = 7, kProbeLdsTab = 8, kProbeCount = 9;
constexpr size_t kStreamBytes = 256ull << 20;
constexpr int kValuChains = 8;
constexpr int kLdsRows = 16;

__global__ __launch_bounds__(64) void probe_valu(uint32_t iters, uint32_t seed, uint32_t *out) {
    const uint32_t lane = threadIdx.x;
    uint32_t x[kValuChains], c[kValuChains];
#pragma unroll
    for (int j = 0; j < kValuChains; ++j) {
        x[j] = (lane + 1) * 0x9E3779B9u ^ (seed + j * 0x85EBCA6Bu) ^ blockIdx.x;
        c[j] = 0;
    }
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t s = i * 0x01000193u + seed;  // wave-uniform (scalar)
#pragma unroll
        for (int j = 0; j < kValuChains; ++j) {
            x[j] ^= s;
            c[j] = __popc(x[j]) + c[j];
        }
    }
    uint32_t t = 0;
#pragma unroll
    for (int j = 0; j < kValuChains; ++j) t += c[j] ^ x[j];
    out[blockIdx.x * 64 + lane] = t;
}

__global__ __launch_bounds__(64) void probe_philox(uint32_t iters, uint32_t seed, uint32_t *out) {
    const uint32_t lane = threadIdx.x;
    uint32_t c[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        c[s][0] = lane + 64 * s;
        c[s][1] = blockIdx.x;
        c[s][2] = seed;
        c[s][3] = kMcTag;
    }
    // each stream's output block is its next input: nothing but the rounds
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int s = 0; s < 4; ++s) philox4x32_10(c[s], seed, seed ^ 0x5bd1e995u);
    }
    uint32_t acc = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) acc += c[s][0] ^ c[s][1] ^ c[s][2] ^ c[s][3];
    out[blockIdx.x * 64 + lane] = acc;
}

// The LDS probes keep kLdsRows rows of 64 counters (4 KB per wave, so the
// resident waves per CU are not limited by LDS); lane l owns column l of
// every row.  Scattered: the lanes of one instruction hit different rows.
__global__ __launch_bounds__(64) void probe_lds(uint32_t iters, uint32_t seed, uint32_t *out) {
    __shared__ uint32_t lds[kLdsRows * kStride];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, kLdsRows);
    uint32_t *p[kLdsRows];
#pragma unroll
    for (int j = 0; j < kLdsRows; ++j) p[j] = lds + ((lane * 7 + j * 5 + seed) % kLdsRows) * kStride + lane;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kLdsRows; ++j) lds_inc(p[j]);
    }
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t r = 0; r < kLdsRows; ++r) t += lds[r * kStride + lane];
    out[blockIdx.x * 64 + lane] = t;
}

// Philox blocks whose words each go through the AN Monte Carlo pair work
// (24-bit multiply, xor, popcount-accumulate): the mix of the Monte Carlo
// kernels without their LDS increments; 40 + 4 x 3 VALU per block.
__global__ __launch_bounds__(64) void probe_mixed(uint32_t iters, uint32_t seed, uint32_t *out) {
    const uint32_t lane = threadIdx.x;
    uint32_t c[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        c[s][0] = lane + 64 * s;
        c[s][1] = blockIdx.x;
        c[s][2] = seed;
        c[s][3] = kMcTag;
    }
    uint32_t w = lane * 0x01000193u + blockIdx.x, acc = 0;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            philox4x32_10(c[s], seed, seed ^ 0x5bd1e995u);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = __popc(w ^ ((c[s][u] & 0x00FFFFFFu) * 61u)) + acc;
        }
    }
    out[blockIdx.x * 64 + lane] = acc ^ c[0][0] ^ c[1][1];
}

// xor + popcount-accumulate on 16 independent chains.
__global__ __launch_bounds__(64) void probe_valu16(uint32_t iters, uint32_t seed, uint32_t *out) {
    const uint32_t lane = threadIdx.x;
    uint32_t x[16], c[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        x[j] = (lane + 1) * 0x9E3779B9u ^ (seed + j * 0x85EBCA6Bu) ^ blockIdx.x;
        c[j] = 0;
    }
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t s = i * 0x01000193u + seed;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            x[j] ^= s;
            c[j] = __popc(x[j]) + c[j];
        }
    }
    uint32_t t = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) t += c[j] ^ x[j];
    out[blockIdx.x * 64 + lane] = t;
}

// LDS increments with every lane of an instruction in the same row (the
// 64 consecutive dwords of one bin), the row changing per instruction.
__global__ __launch_bounds__(64) void probe_lds_row(uint32_t iters, uint32_t seed, uint32_t *out) {
    __shared__ uint32_t lds[kLdsRows * kStride];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, kLdsRows);
    uint32_t *p[kLdsRows];
#pragma unroll
    for (int j = 0; j < kLdsRows; ++j) p[j] = lds + ((j * 5 + seed) % kLdsRows) * kStride + lane;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kLdsRows; ++j) lds_inc(p[j]);
    }
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t r = 0; r < kLdsRows; ++r) t += lds[r * kStride + lane];
    out[blockIdx.x * 64 + lane] = t;
}

// Streaming stores / loads: `iters` passes over n uint4 with a grid stride.
// Byte lookups in the Hamming syndrome tables' layout: 8 tables of 128 bytes
// (32 dwords each, one per LDS bank), read at one per-lane index on 4
// independent chains; the next index depends on the 8 bytes read (three
// v_bitop3_b32 and one multiply-xor per 8 reads, so the LDS and not the VALU
// sets the rate).
constexpr int kTabChains = 4;
__global__ __launch_bounds__(64) void probe_lds_tab(uint32_t iters, uint32_t seed, uint32_t *out) {
    __shared__ uint32_t tab[8 * 32];
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 8 * 32; i += 64) tab[i] = (i + seed) * 2654435761u;
    __syncthreads();
    const uint8_t *tb = reinterpret_cast<const uint8_t *>(tab);
    uint32_t x[kTabChains];
#pragma unroll
    for (int j = 0; j < kTabChains; ++j) x[j] = (lane * 7919u + j * 104729u + seed + blockIdx.x) * 2654435761u;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < kTabChains; ++j) {
            const uint32_t idx = x[j] & 127;
            uint32_t v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) v[t] = tb[128 * t + idx];
            uint32_t acc = __builtin_amdgcn_bitop3_b32(v[0], v[1], v[2], 0x96);
            acc = __builtin_amdgcn_bitop3_b32(acc, v[3], v[4], 0x96);
            acc = __builtin_amdgcn_bitop3_b32(acc, v[5], v[6], 0x96);
            x[j] = __builtin_amdgcn_bitop3_b32(acc, v[7], x[j] * 0x01000193u, 0x96);
        }
    }
    uint32_t a = 0;
#pragma unroll
    for (int j = 0; j < kTabChains; ++j) a ^= x[j];
    out[blockIdx.x * 64 + lane] = a;
}

__global__ __launch_bounds__(256) void probe_write(uint32_t iters, uint32_t seed, uint4 *buf, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint4 x = make_uint4(seed, threadIdx.x, blockIdx.x, 0);
    for (uint32_t it = 0; it < iters; ++it)
        for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) buf[e] = x;
}
__global__ __launch_bounds__(256) void probe_read(uint32_t iters, uint32_t seed, const uint4 *buf, uint64_t n,
                                                  uint32_t *out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t acc = seed;
    for (uint32_t it = 0; it < iters; ++it)
        for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
            const uint4 v = buf[e];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    if (acc == 0x9E3779B9u) out[0] = acc;  // keeps the loads
}

}  // namespace
}  // namespace cr

using namespace cr;

extern "C" cr_status cr_probe_ceiling(cr_context *ctx, uint32_t probe, uint32_t iters, double *units_per_s,
                                      double *kernel_ms) {
    clear_error();
    if (!ctx || !units_per_s) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (probe >= kProbeCount) return fail(CR_ERR_INVALID_ARGUMENT, "unknown probe %u", probe);
    if (iters == 0 || iters > (1u << 24)) return fail(CR_ERR_INVALID_ARGUMENT, "iters=%u outside [1, 2^24]", iters);
    CR_HIP(hipSetDevice(ctx->device));
    if (probe == kProbeWrite || probe == kProbeRead) {
        if (iters > 64) return fail(CR_ERR_INVALID_ARGUMENT, "streaming probes: iters=%u above 64 passes", iters);
        CR_TRY(ctx->table_b.reserve(kStreamBytes + 256));
        uint4 *buf = ctx->table_b.as<uint4>();
        const uint64_t n = kStreamBytes / sizeof(uint4);
        const uint32_t blocks = (uint32_t)ctx->cus * 8;
        CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        if (probe == kProbeWrite)
            hipLaunchKernelGGL(probe_write, dim3(blocks), dim3(256), 0, ctx->stream, iters, 0x2545F491u, buf, n);
        else
            hipLaunchKernelGGL(probe_read, dim3(blocks), dim3(256), 0, ctx->stream, iters, 0x2545F491u,
                               (const uint4 *)buf, n, reinterpret_cast<uint32_t *>(buf + n));
        CR_HIP(hipGetLastError());
        CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        CR_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *units_per_s = ms > 0 ? (double)kStreamBytes * iters / (ms * 1e-3) : 0;
        if (kernel_ms) *kernel_ms = ms;
        return CR_OK;
    }
    const void *kernels[kProbeCount] = {
        reinterpret_cast<const void *>(&probe_valu),   reinterpret_cast<const void *>(&probe_philox),
        reinterpret_cast<const void *>(&probe_lds),    reinterpret_cast<const void *>(&probe_mixed),
        reinterpret_cast<const void *>(&probe_valu16), reinterpret_cast<const void *>(&probe_lds_row),
        nullptr,                                       nullptr,
        reinterpret_cast<const void *>(&probe_lds_tab)};
    const void *k = kernels[probe];
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess || per_cu <= 0)
        return fail(CR_ERR_DEVICE, "occupancy query failed for probe %u", probe);
    const uint32_t waves = (uint32_t)per_cu * (uint32_t)ctx->cus;
    CR_TRY(ctx->table_b.reserve((size_t)waves * 64 * sizeof(uint32_t)));
    uint32_t a_iters = iters, a_seed = 0x2545F491u;
    uint32_t *a_out = ctx->table_b.as<uint32_t>();
    void *args[] = {&a_iters, &a_seed, &a_out};
    CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    CR_HIP(hipLaunchKernel(k, dim3(waves), dim3(64), args, 0, ctx->stream));
    CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    // units per lane and loop iteration (VALU instructions of the loop body,
    // LDS increments or LDS reads)
    const double per_lane_iter[kProbeCount] = {2.0 * kValuChains, 4 * 40.0, (double)kLdsRows,
                                               2 * (40.0 + 4 * 3), 32.0, (double)kLdsRows, 0.0, 0.0,
                                               8.0 * kTabChains};
    const double units = (double)waves * 64.0 * (double)iters * per_lane_iter[probe];
    *units_per_s = ms > 0 ? units / (ms * 1e-3) : 0;
    if (kernel_ms) *kernel_ms = ms;
    return CR_OK;
}
