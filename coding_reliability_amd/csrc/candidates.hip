// ANcandidates on the GPU (gfx950): for every odd A with exactly `abits`
// bits, the minimum over d in [1, 2^k) of popcount(c ^ 3c), c = A*d -- the
// length of the shortest signed-digit (NAF) representation of c -- and the
// maximum of these minima over A (worker, ANcandidates/ANcandidates.cpp:35-160).
//
// The kernel scans a d range for a list of A (one lane per A): work item =
// (group of 64 list entries, chunk of the d range); the lane walks
// c = A*d0, A*(d0+1), ... in registers (c += A, 3c by a shift-add, two-dword
// xor + bcnt, min).  Chunk minima meet in vals[i] (atomic min); done[i] counts
// finished chunks and the chunk that completes entry i publishes its minimum
// into the proven best `gbest` (atomic max).  A chunk is skipped for an entry
// whose minimum so far is already below gbest, and a wave stops once every
// lane's running minimum is below gbest (checked every kCheck steps).
//
// Host schedule (three launches), so that waves carry only A's that can still
// attain the maximum (the SIMD lanes of a wave run in lock step, and with
// about a quarter of all A attaining the maximum every fixed group of 64 A
// holds some, so per-wave pruning alone saves little):
//   1. prefix: all A over d < P (P = 4096): prefix minima pm[A] >= min(A)
//      (no pruning: a prefix minimum proves nothing about the maximum);
//   2. bound: the 64 A with the largest pm over the rest of the range: their
//      exact minima give L, a minimum that some A attains, so max >= L;
//   3. the A whose minimum so far is >= L (the others cannot attain the
//      maximum) over the rest of the range in rounds of 8x growing d ranges,
//      compacted into full waves before each round, with gbest at L.
// Every A attaining the maximum survives each step and is scanned in full, so
// the maximum and the set of A attaining it are exact.
#include <algorithm>
#include <numeric>
#include <vector>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

constexpr uint32_t kUnrollC = 8;
constexpr uint64_t kCheck = 1u << 11;      // steps between early-exit checks
constexpr uint64_t kMinChunk = 1u << 14;   // smallest d chunk worth an item
constexpr uint64_t kPrefix = 1u << 12;     // d range of the prefix pass
constexpr uint64_t kBoundA = 64;           // A scanned in full for the lower bound
constexpr uint64_t kRoundGrowth = 8;       // d range growth per survivor round

__global__ void cand_init_kernel(uint32_t *__restrict__ done, unsigned long long *__restrict__ ctr, uint64_t n,
                                 uint32_t gbest0) {
    const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < n; i += (uint64_t)gridDim.x * blockDim.x) done[i] = 0;
    if (i0 == 0) {
        ctr[0] = 0;       // work-queue head
        ctr[1] = gbest0;  // proven best (low dword)
    }
}

__global__ __launch_bounds__(64) void cand_kernel(const uint64_t *__restrict__ alist, uint64_t minA, uint64_t numA,
                                                  uint64_t dstart, uint64_t dend, uint64_t chunk, uint32_t nchunks,
                                                  uint64_t total_items, uint32_t publish,
                                                  uint32_t *__restrict__ vals, uint32_t *__restrict__ done,
                                                  unsigned long long *__restrict__ queue,
                                                  unsigned int *__restrict__ gbest) {
    const uint32_t lane = threadIdx.x;
    for (uint64_t t = first_item(); t < total_items; t = following_item(queue, t)) {
        const uint64_t g = uniform64(t / nchunks);
        const uint32_t ci = uniform32((uint32_t)(t - g * nchunks));
        const uint64_t idx = g * 64 + lane;
        const bool valid = idx < numA;
        const uint64_t A = valid ? (alist ? alist[idx] : minA + 2 * idx) : minA;
        const uint64_t d0 = dstart + (uint64_t)ci * chunk;
        const uint64_t d1 = d0 + chunk < dend ? d0 + chunk : dend;
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
            if (publish && old == nchunks - 1)
                atomicMax(gbest, __hip_atomic_load(vals + idx, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
        }
    }
}

}  // namespace

// One scan: vals (n entries, host in/out: running minima) over d in [dstart,
// dend) for the A of `alist` (host, n entries) or of minA + 2i when alist is
// empty; gbest starts at gbest0.  publish: a completed entry's value is its
// exact minimum over all d (true once the scan ends at 2^k and vals came in
// with the minima below dstart), so it may raise gbest; otherwise gbest stays
// at gbest0.  Adds the kernel time to *ms.
static cr_status cand_scan(cr_context *ctx, const std::vector<uint64_t> &alist, uint64_t minA, uint64_t n,
                           uint64_t dstart, uint64_t dend, uint32_t gbest0, bool publish, std::vector<uint32_t> &vals,
                           double *ms) {
    if (n == 0 || dend <= dstart) return CR_OK;
    const uint64_t steps = dend - dstart, ngroups = (n + 63) / 64;
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
    CR_TRY(ctx->table_a.reserve(n * sizeof(uint32_t)));
    CR_TRY(ctx->table_b.reserve(n * sizeof(uint32_t)));
    CR_TRY(ctx->counters.reserve(2 * sizeof(uint64_t)));
    uint32_t *d_vals = ctx->table_a.as<uint32_t>(), *d_done = ctx->table_b.as<uint32_t>();
    uint64_t *d_list = nullptr;
    if (!alist.empty()) {
        CR_TRY(ctx->jobs.reserve(n * sizeof(uint64_t)));
        d_list = ctx->jobs.as<uint64_t>();
        CR_HIP(hipMemcpyAsync(d_list, alist.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    }
    CR_HIP(hipMemcpyAsync(d_vals, vals.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    unsigned long long *ctr = ctx->counters.as<unsigned long long>();
    const uint32_t iblocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(cand_init_kernel, dim3(iblocks), dim3(256), 0, ctx->stream, d_done, ctr, n, gbest0);
    CR_HIP(hipGetLastError());
    const uint64_t blocks = std::min<uint64_t>(waves, total_items);
    CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(cand_kernel, dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, d_list, minA, n, dstart, dend,
                       chunk, (uint32_t)nchunks, total_items, (uint32_t)publish, d_vals, d_done, ctr,
                       reinterpret_cast<unsigned int *>(ctr + 1));
    CR_HIP(hipGetLastError());
    CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    CR_HIP(hipMemcpyAsync(vals.data(), d_vals, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    float t = 0;
    CR_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms += t;
    return CR_OK;
}

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
    const uint64_t nwords = 1ull << k;  // d = 1 .. 2^k - 1
    const uint64_t P = std::min<uint64_t>(nwords, kPrefix);
    double ms = 0;
    // 1. prefix minima of every A
    std::vector<uint32_t> v(numA, 0xFFFFFFFFu);
    CR_TRY(cand_scan(ctx, {}, minA, numA, 1, P, 0, P == nwords, v, &ms));
    if (P < nwords) {
        // 2. exact minima of the A with the largest prefix minima: lower bound L
        std::vector<uint64_t> order(numA);
        std::iota(order.begin(), order.end(), 0);
        const uint64_t nb = std::min<uint64_t>(numA, kBoundA);
        std::partial_sort(order.begin(), order.begin() + nb, order.end(),
                          [&](uint64_t a, uint64_t b) { return v[a] != v[b] ? v[a] > v[b] : a < b; });
        std::vector<uint64_t> la(nb);
        std::vector<uint32_t> lv(nb);
        for (uint64_t i = 0; i < nb; ++i) {
            la[i] = minA + 2 * order[i];
            lv[i] = v[order[i]];
        }
        CR_TRY(cand_scan(ctx, la, 0, nb, P, nwords, 0, true, lv, &ms));
        uint32_t L = 0;
        for (uint64_t i = 0; i < nb; ++i) {
            v[order[i]] = lv[i];
            L = std::max(L, lv[i]);
        }
        // 3. the remaining A whose minimum so far reaches L, over d ranges
        //    growing 8x per round, recompacted between rounds
        std::vector<uint8_t> in_bound(numA, 0);
        for (uint64_t i = 0; i < nb; ++i) in_bound[order[i]] = 1;
        for (uint64_t lo = P; lo < nwords;) {
            const uint64_t hi = lo > nwords / kRoundGrowth ? nwords : lo * kRoundGrowth;
            std::vector<uint64_t> sa, si;
            std::vector<uint32_t> sv;
            for (uint64_t i = 0; i < numA; ++i)
                if (!in_bound[i] && v[i] >= L) {
                    si.push_back(i);
                    sa.push_back(minA + 2 * i);
                    sv.push_back(v[i]);
                }
            if (si.empty()) break;
            CR_TRY(cand_scan(ctx, sa, 0, sa.size(), lo, hi, L, hi == nwords, sv, &ms));
            for (size_t j = 0; j < si.size(); ++j) v[si[j]] = sv[j];
            lo = hi;
        }
    }
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
    if (kernel_ms) *kernel_ms = ms;
    return CR_OK;
}
