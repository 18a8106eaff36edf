// ANcandidates on the GPU (gfx950): for every odd A with exactly `abits`
// bits, the minimum over d in [1, 2^k) of popcount(c ^ 3c), c = A*d -- the
// length of the shortest signed-digit (NAF) representation of c -- and the
// maximum of these minima over A (worker, ANcandidates/ANcandidates.cpp:35-160).
//
// Work item = (group of 64 consecutive A, chunk of the d range): one lane per
// A walks c = A*d0, A*(d0+1), ... in registers (c += A, 3c by a shift-add,
// two-dword xor + bcnt, min).  Items are taken group-major from a queue, so
// the chunks of the first groups run side by side and their A's complete
// early.  Per A, chunk minima meet in vals[A] (atomic min) and done[A] counts
// finished chunks; the chunk that completes an A publishes its minimum into
// the proven best `gbest` (atomic max).  Pruning mirrors the reference's
// break: a chunk is skipped for an A whose partial minimum is already below
// gbest, and a wave stops once every lane's running minimum is below gbest
// (checked every kCheck steps).  Such A's cannot attain the maximum (their
// minimum is below a minimum some A attains), every A that does is never
// pruned, so the maximum and the set of A attaining it are exact.
#include <algorithm>
#include <vector>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kUnrollC = 8;
constexpr uint64_t kCheck = 1u << 11;      // steps between early-exit checks
constexpr uint64_t kMinChunk = 1u << 14;   // smallest d chunk worth an item

__global__ void cand_init_kernel(uint32_t *__restrict__ vals, uint32_t *__restrict__ done, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        vals[i] = 0xFFFFFFFFu;
        done[i] = 0;
    }
}

__global__ __launch_bounds__(64) void cand_kernel(uint64_t minA, uint64_t numA, uint64_t nwords, uint64_t chunk,
                                                  uint32_t nchunks, uint64_t total_items,
                                                  uint32_t *__restrict__ vals, uint32_t *__restrict__ done,
                                                  unsigned long long *__restrict__ queue,
                                                  unsigned int *__restrict__ gbest) {
    const uint32_t lane = threadIdx.x;
    for (;;) {
        const uint64_t t = next_item(queue);
        if (t >= total_items) break;
        const uint64_t g = uniform64(t / nchunks);
        const uint32_t ci = uniform32((uint32_t)(t - g * nchunks));
        const uint64_t idx = g * 64 + lane;
        const bool valid = idx < numA;
        const uint64_t A = minA + 2 * (valid ? idx : 0);
        const uint64_t d0 = 1 + (uint64_t)ci * chunk;
        const uint64_t d1 = d0 + chunk < nwords ? d0 + chunk : nwords;
        uint32_t best = __hip_atomic_load(gbest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool active =
            valid && __hip_atomic_load(vals + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= best;
        uint32_t mn = 0xFFFFFFFFu;
        if (__any(active)) {
            uint64_t c = A * d0;
            uint64_t s = d0;
            while (s < d1) {
                const uint64_t lim = s + kCheck < d1 ? s + kCheck : d1;
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
                best = __hip_atomic_load(gbest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (s < d1 && __all(!active || mn < best)) break;
            }
        }
        if (valid) {
            if (active) atomicMin(vals + idx, mn);
            const uint32_t old = __hip_atomic_fetch_add(done + idx, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == nchunks - 1)
                atomicMax(gbest, __hip_atomic_load(vals + idx, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
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
    const uint64_t nwords = 1ull << k, steps = nwords - 1;  // d = 1 .. 2^k - 1
    const uint64_t ngroups = (numA + 63) / 64;
    int per_cu = 0;
    CR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&cand_kernel), 64, 0));
    const uint64_t waves = (uint64_t)std::max(per_cu, 1) * ctx->cus;
    // about 8 items per resident wave, chunks of at least kMinChunk steps
    uint64_t nchunks = (8 * waves + ngroups - 1) / ngroups;
    nchunks = std::min<uint64_t>(nchunks, (steps + kMinChunk - 1) / kMinChunk);
    nchunks = std::max<uint64_t>(nchunks, 1);
    const uint64_t chunk = (steps + nchunks - 1) / nchunks;
    nchunks = (steps + chunk - 1) / chunk;
    const uint64_t total_items = ngroups * nchunks;
    CR_TRY(ctx->table_a.reserve(numA * sizeof(uint32_t)));
    CR_TRY(ctx->table_b.reserve(numA * sizeof(uint32_t)));
    CR_TRY(ctx->counters.reserve(2 * sizeof(uint64_t)));
    uint32_t *vals = ctx->table_a.as<uint32_t>(), *done = ctx->table_b.as<uint32_t>();
    unsigned long long *ctr = ctx->counters.as<unsigned long long>();
    CR_TRY(fill_zero_u64(ctr, 2, ctx->stream));
    const uint32_t iblocks = (uint32_t)std::min<uint64_t>((numA + 255) / 256, 4096);
    hipLaunchKernelGGL(cand_init_kernel, dim3(iblocks), dim3(256), 0, ctx->stream, vals, done, numA);
    CR_HIP(hipGetLastError());
    const uint64_t blocks = std::min<uint64_t>(waves, total_items);
    CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(cand_kernel, dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, minA, numA, nwords, chunk,
                       (uint32_t)nchunks, total_items, vals, done, ctr, reinterpret_cast<unsigned int *>(ctr + 1));
    CR_HIP(hipGetLastError());
    CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<uint32_t> v(numA);
    CR_HIP(hipMemcpyAsync(v.data(), vals, numA * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t b = 0;
    for (uint32_t x : v) b = x > b ? x : b;
    *best = b;
    uint64_t n = 0;
    for (uint64_t i = 0; i < numA; ++i)
        if (v[i] == b) {
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
