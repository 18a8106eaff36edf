// ANcandidates on the GPU (gfx950): for every odd A with exactly `abits`
// bits, the minimum over d in [1, 2^k) of popcount(c ^ 3c), c = A*d -- the
// length of the shortest signed-digit (NAF) representation of c -- and the
// maximum of these minima over A (worker, ANcandidates/ANcandidates.cpp:35-160).
//
// One lane per A: the lane walks c = A, 2A, ... in registers (c += A, 3c by a
// shift-add, two-dword xor + bcnt, min).  Early exit mirrors the reference's
// break: a wave stops once every lane's running minimum is below the best
// minimum already proven by a completed A (read from device memory every
// kCheck steps); such A's cannot attain the maximum, and every A that does
// runs to completion, so the maximum and the set of A attaining it are exact.
#include <algorithm>
#include <vector>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kUnrollC = 8;
constexpr uint64_t kCheck = 1u << 14;  // steps between early-exit checks

__global__ __launch_bounds__(64) void cand_kernel(uint64_t minA, uint64_t numA, uint32_t k, uint32_t *__restrict__ vals,
                                                  unsigned int *__restrict__ gbest) {
    const uint64_t ngroups = (numA + 63) / 64;
    const uint64_t steps = (1ull << k) - 1;  // d = 1 .. 2^k - 1
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t idx = g * 64 + threadIdx.x;
        const bool valid = idx < numA;
        const uint64_t A = minA + 2 * (valid ? idx : 0);
        uint64_t c = A;
        uint32_t mn = 0xFFFFFFFFu;
        bool pruned = false;
        uint64_t s = 0;
        while (s < steps) {
            uint64_t lim = s + kCheck;
            if (lim > steps) lim = steps;
            for (; s + kUnrollC <= lim; s += kUnrollC) {
#pragma unroll
                for (uint32_t u = 0; u < kUnrollC; ++u) {
                    const uint32_t p = (uint32_t)__popcll(c ^ ((c << 1) + c));
                    mn = p < mn ? p : mn;
                    c += A;
                }
            }
            for (; s < lim; ++s) {
                const uint32_t p = (uint32_t)__popcll(c ^ ((c << 1) + c));
                mn = p < mn ? p : mn;
                c += A;
            }
            const uint32_t best = __hip_atomic_load(gbest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (s < steps && __all(!valid || mn < best)) {
                pruned = true;
                break;
            }
        }
        if (valid) {
            vals[idx] = mn;
            if (!pruned) atomicMax(gbest, mn);
        }
    }
}

}  // namespace

}  // namespace cr

using namespace cr;

extern "C" cr_status cr_an_candidates(cr_context *ctx, uint32_t k, uint32_t abits, uint64_t *best, uint64_t *cands,
                                      uint64_t cap, uint64_t *ncands, double *kernel_ms) {
    clear_error();
    if (!ctx || !best || !ncands) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    *best = 0;
    *ncands = 0;
    if (k < 1 || abits < 2) return CR_OK;  // the reference's loops are empty there
    if (k + abits > 62) return fail(CR_ERR_UNSUPPORTED, "k + |A| = %u exceeds 62 bits", k + abits);
    if (abits > 40) return fail(CR_ERR_UNSUPPORTED, "|A| = %u", abits);
    CR_HIP(hipSetDevice(ctx->device));
    const uint64_t minA = (1ull << (abits - 1)) + 1, maxA = (1ull << abits) - 1;
    const uint64_t numA = (maxA - minA) / 2 + 1;
    CR_TRY(ctx->table_a.reserve(numA * sizeof(uint32_t)));
    CR_TRY(ctx->counters.reserve(sizeof(uint64_t)));
    CR_TRY(fill_zero_u64(ctx->counters.as<unsigned long long>(), 1, ctx->stream));
    const uint64_t ngroups = (numA + 63) / 64;
    int per_cu = 0;
    CR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&cand_kernel), 64, 0));
    uint64_t blocks = (uint64_t)std::max(per_cu, 1) * ctx->cus;
    if (blocks > ngroups) blocks = ngroups;
    CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(cand_kernel, dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, minA, numA, k,
                       ctx->table_a.as<uint32_t>(), reinterpret_cast<unsigned int *>(ctx->counters.ptr));
    CR_HIP(hipGetLastError());
    CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<uint32_t> vals(numA);
    CR_HIP(hipMemcpyAsync(vals.data(), ctx->table_a.ptr, numA * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t b = 0;
    for (uint32_t v : vals) b = v > b ? v : b;
    *best = b;
    uint64_t n = 0;
    for (uint64_t i = 0; i < numA; ++i)
        if (vals[i] == b) {
            if (cands && n < cap) cands[n] = minA + 2 * i;
            ++n;
        }
    *ncands = n;
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}
