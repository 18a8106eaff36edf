// Issue-rate probes: the device ceilings the histogram kernels are bound by,
// measured on the device they run on (gfx950, wave64).  Each probe runs the
// instruction mix of one kind of kernel and nothing else, on every resident
// single-wave workgroup, for a fixed number of iterations:
//   kProbeValu   xor + popcount-accumulate (v_xor_b32, v_bcnt_u32_b32) on 8
//                independent chains: the pair, grid and exhaustive Hamming
//                inner loops; 2 lane-ops per chain step
//   kProbePhilox Philox4x32-10 blocks (4 independent streams per lane): the
//                Monte Carlo kernels' generator; 60 lane-ops per block
//                (10 rounds of 2 multiplies, 2 high halves and 2 xors), the
//                count the rooflines use for it
//   kProbeLds    wave-wide ds_add_u32 to lane-private LDS counters (row
//                stride 64 dwords, the histogram layout of every kernel):
//                the per-pair histogram increment; 1 unit per lane increment
#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kProbeValu = 0, kProbePhilox = 1, kProbeLds = 2;
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

__global__ __launch_bounds__(64) void probe_lds(uint32_t iters, uint32_t seed, uint32_t *out) {
    __shared__ uint32_t lds[64 * kStride];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, 64);
    // kLdsRows different rows per lane (fixed over the loop), each in the
    // lane's own bank
    uint32_t *p[kLdsRows];
#pragma unroll
    for (int j = 0; j < kLdsRows; ++j) p[j] = lds + ((lane * 7 + j * 5 + seed) & 63) * kStride + lane;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kLdsRows; ++j) lds_inc(p[j]);
    }
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t r = 0; r < 64; ++r) t += lds[r * kStride + lane];
    out[blockIdx.x * 64 + lane] = t;
}

}  // namespace
}  // namespace cr

using namespace cr;

extern "C" cr_status cr_probe_ceiling(cr_context *ctx, uint32_t probe, uint32_t iters, double *units_per_s,
                                      double *kernel_ms) {
    clear_error();
    if (!ctx || !units_per_s) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (probe > kProbeLds) return fail(CR_ERR_INVALID_ARGUMENT, "unknown probe %u", probe);
    if (iters == 0 || iters > (1u << 24)) return fail(CR_ERR_INVALID_ARGUMENT, "iters=%u outside [1, 2^24]", iters);
    CR_HIP(hipSetDevice(ctx->device));
    const void *k = probe == kProbeValu     ? reinterpret_cast<const void *>(&probe_valu)
                    : probe == kProbePhilox ? reinterpret_cast<const void *>(&probe_philox)
                                            : reinterpret_cast<const void *>(&probe_lds);
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
    const double per_lane_iter = probe == kProbeValu ? 2.0 * kValuChains : probe == kProbePhilox ? 4 * 60.0 : kLdsRows;
    const double units = (double)waves * 64.0 * (double)iters * per_lane_iter;
    *units_per_s = ms > 0 ? units / (ms * 1e-3) : 0;
    if (kernel_ms) *kernel_ms = ms;
    return CR_OK;
}
