// Host side of the Extended Hamming and AN Monte Carlo entry points.
#include <algorithm>
#include <cstring>
#include <vector>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {

// The reference's (39,32) parity masks (distance_distribution/src/hamming.h:54-64)
// are the low halves of the (72,64) patterns of legacy/Hamming/ExtHamming64.hpp:31-37;
// one table truncated to the data width serves every k.
static const uint64_t kHamPatterns[7] = {
    0xAB55555556AAAD5Bull, 0xCD9999999B33366Dull, 0x78F1E1E1E3C3C78Eull, 0x01FE01FE03FC07F0ull,
    0x01FFFE0003FFF800ull, 0x01FFFFFFFC000000ull, 0xFE00000000000000ull};

uint32_t hamming_masks(uint32_t k, uint64_t masks[7]) {
    const uint64_t dm = k >= 64 ? ~0ull : ((1ull << k) - 1);
    uint32_t r = 0;
    for (int t = 0; t < 7; ++t) {
        masks[t] = kHamPatterns[t] & dm;
        if (masks[t]) ++r;
    }
    return r;
}

}  // namespace cr

using namespace cr;

extern "C" uint32_t cr_hamming_parity_bits(uint32_t k) {
    uint64_t m[7];
    return hamming_masks(k, m);
}

namespace {

// Requests [0, n) as one Hamming job each: the jobs of 32-bit words in one
// launch, those of 64-bit words (k > 32) in a second on the same stream, each
// launch's jobs in request order; requests without samples launch nothing.
// hists[i * nbins + b] is request i's count of weight b.
cr_status hamming_run(cr_context *ctx, const cr_ham_request *reqs, uint32_t n, uint64_t *hists, uint32_t nbins,
                      double *kernel_ms) {
    if (!ctx || !reqs || !hists) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || n > kHamMaxBatch)
        return fail(CR_ERR_INVALID_ARGUMENT, "%u requests outside [1, %u]", n, kHamMaxBatch);
    struct Launch {
        std::vector<HamJob> jobs;
        std::vector<uint32_t> index, needs;  // request of each job, and its weight bins
        std::vector<double> rates;
        double est_s = 0;
        uint32_t rows = 2;
        uint64_t items = 0;
        bool short_launch = false;
    } launches[2];
    uint64_t list_words = 0;
    for (uint32_t q = 0; q < n; ++q) {
        const cr_ham_request *req = reqs + q;
        if (req->k < 1 || req->k > 64) return fail(CR_ERR_INVALID_ARGUMENT, "k=%u outside [1, 64]", req->k);
        if (req->source > CR_HAM_MC) return fail(CR_ERR_INVALID_ARGUMENT, "unknown source %u", req->source);
        if (req->source == CR_HAM_LIST && req->count && !req->list)
            return fail(CR_ERR_INVALID_ARGUMENT, "list source without a list");
        const uint32_t per = req->k <= 32 ? 4 : 2;
        if (req->source == CR_HAM_MC && req->first % per)
            return fail(CR_ERR_INVALID_ARGUMENT, "Monte Carlo first sample must be a multiple of %u", per);
        HamJob job{};
        const uint32_t r = hamming_masks(req->k, job.masks);
        const uint32_t need = req->k + r + 1;  // weights 0..k+r
        if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but weights reach %u", nbins, need - 1);
        if (req->count == 0) continue;
        job.first = req->first;
        job.stride = req->stride;
        job.count = req->count;
        job.seed = req->seed;
        job.source = req->source;
        job.k = req->k;
        job.stream = req->stream;
        job.nbins = need;
        // Measured rates (words/s): exhaustive 7.5e12, 32-bit range/stride
        // 1.9e12, 64-bit 0.55e12, Monte Carlo 1.55e12 / 0.75e12, lists about
        // half the stride rates.
        double rate = req->k > 32 ? 0.55e12 : 1.9e12;
        if (req->source == CR_HAM_RANGE && req->first == 0) rate = 7.5e12;
        if (req->source == CR_HAM_MC) rate = req->k > 32 ? 0.75e12 : 1.55e12;
        if (req->source == CR_HAM_LIST) {
            rate *= 0.5;
            list_words += req->count;
        }
        Launch &L = launches[req->k > 32];
        L.est_s += (double)req->count / rate;
        L.rows = std::max(L.rows, need + 1);
        L.jobs.push_back(job);
        L.index.push_back(q);
        L.needs.push_back(need);
        L.rates.push_back(rate);
    }
    CR_HIP(hipSetDevice(ctx->device));
    memset(hists, 0, (size_t)n * nbins * sizeof(uint64_t));
    if (kernel_ms) *kernel_ms = 0;
    if (launches[0].jobs.empty() && launches[1].jobs.empty()) return CR_OK;
    if (list_words) CR_TRY(ctx->table_a.reserve(list_words * sizeof(uint64_t)));
    uint64_t list_at = 0, entries = 0;
    std::vector<uint64_t> out_at[2];
    for (Launch &L : launches) {
        // Work assignment.  The queue balances waves that advance at different
        // rates (a static assignment measured about 25 % slower on launches
        // of 0.5 ms and more), but each queue step is a device-scope atomic
        // on one counter, serialized at about 12.5 ns, one per item plus one
        // per wave: 40-150 us per launch.  Launches estimated below 100 us of
        // work therefore run with a static assignment.
        L.short_launch = L.est_s < 100e-6;
        const uint64_t waves = (uint64_t)hamming_waves_per_cu(L.rows) * ctx->cus;
        std::vector<uint64_t> &at = out_at[&L - launches];
        at.resize(L.jobs.size());
        for (size_t j = 0; j < L.jobs.size(); ++j) {
            HamJob &job = L.jobs[j];
            // items of 64 x spl words, spl the smallest power of two in [16,
            // kHamMaxLaneSamples] that gives at most 8 items per resident
            // wave for a static assignment or 32 for the queue (the last
            // items set the launch's tail: with at most 4 per wave a
            // 2^30-word (72,64) Monte Carlo launch ran 20 % below the rate of
            // 2^34 words), and for the queue at least rate/8e7 words per item
            // (the counter's 8e7 items/s)
            const uint64_t per_wave = L.short_launch ? 8 : 32;
            const uint64_t want = (job.count + 64 * per_wave * waves - 1) / (64 * per_wave * waves);
            const double min_words = L.short_launch ? 0.0 : L.rates[j] / 8e7;
            uint64_t spl = 16;
            while ((spl < want || 64.0 * (double)spl < min_words) && spl < kHamMaxLaneSamples) spl *= 2;
            job.lane_samples = (uint32_t)spl;
            job.items = hamming_items(job.count, job.lane_samples);
            job.item_begin = L.items;
            L.items += job.items;
            if (job.source == kHamList) {
                const cr_ham_request *req = reqs + L.index[j];
                CR_HIP(hipMemcpyAsync(ctx->table_a.as<uint64_t>() + list_at, req->list,
                                      req->count * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
                job.list = ctx->table_a.as<uint64_t>() + list_at;
                list_at += req->count;
            }
            // dev::kHistCopies line-aligned copies of each histogram, summed
            // on the host
            job.out_copies = dev::kHistCopies;
            at[j] = entries;
            entries += dev::hist_copy_stride(L.needs[j]) * job.out_copies;
        }
    }
    const size_t njobs = launches[0].jobs.size() + launches[1].jobs.size();
    CR_TRY(ctx->hist.reserve(entries * sizeof(uint64_t)));
    CR_TRY(ctx->counters.reserve(2 * sizeof(uint64_t)));
    CR_TRY(ctx->jobs.reserve(njobs * sizeof(HamJob)));
    std::vector<HamJob> all;
    all.reserve(njobs);
    for (Launch &L : launches)
        for (size_t j = 0; j < L.jobs.size(); ++j) {
            L.jobs[j].out = ctx->hist.as<unsigned long long>() + out_at[&L - launches][j];
            all.push_back(L.jobs[j]);
        }
    CR_TRY(fill_zero_u64(ctx->hist.as<unsigned long long>(), entries, ctx->stream));
    CR_HIP(hipMemcpyAsync(ctx->jobs.ptr, all.data(), njobs * sizeof(HamJob), hipMemcpyHostToDevice, ctx->stream));
    // HIP events before the first launch and after the last
    const bool both = !launches[0].jobs.empty() && !launches[1].jobs.empty();
    for (int w = 0, first = 1; w < 2; ++w) {
        const Launch &L = launches[w];
        if (L.jobs.empty()) continue;
        const bool last = w == 1 || !both;
        CR_TRY(launch_hamming(ctx->jobs.as<HamJob>() + (w ? launches[0].jobs.size() : 0), (uint32_t)L.jobs.size(),
                              L.items, w == 1, ctx->counters.as<unsigned long long>() + w, L.short_launch, L.rows,
                              ctx->cus, ctx->stream, kernel_ms && first ? ctx->ev0 : nullptr,
                              kernel_ms && last ? ctx->ev1 : nullptr));
        first = 0;
    }
    CR_TRY(ctx->host.reserve(entries * sizeof(uint64_t)));
    CR_HIP(hipMemcpyAsync(ctx->host.ptr, ctx->hist.ptr, entries * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t *copies = static_cast<const uint64_t *>(ctx->host.ptr);
    for (const Launch &L : launches)
        for (size_t j = 0; j < L.jobs.size(); ++j) {
            const uint64_t cstride = dev::hist_copy_stride(L.needs[j]);
            const uint64_t base = out_at[&L - launches][j];
            uint64_t *hist = hists + (uint64_t)L.index[j] * nbins;
            for (uint32_t b = 0; b < L.needs[j]; ++b) {
                uint64_t sum = 0;
                for (uint32_t c = 0; c < L.jobs[j].out_copies; ++c) sum += copies[base + c * cstride + b];
                hist[b] = sum;
            }
        }
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}

}  // namespace

extern "C" cr_status cr_hamming_weights(cr_context *ctx, const cr_ham_request *req, uint64_t *hist, uint32_t nbins,
                                        double *kernel_ms) {
    clear_error();
    return hamming_run(ctx, req, 1, hist, nbins, kernel_ms);
}

extern "C" cr_status cr_hamming_weights_batch(cr_context *ctx, const cr_ham_request *reqs, uint32_t n,
                                              uint64_t *hists, uint32_t nbins, double *kernel_ms) {
    clear_error();
    return hamming_run(ctx, reqs, n, hists, nbins, kernel_ms);
}

extern "C" cr_status cr_hamming_report(uint32_t k, const uint64_t *weights, int with_1bit, uint64_t *out_u128) {
    clear_error();
    if (!weights || !out_u128) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (k < 1 || k > 64) return fail(CR_ERR_INVALID_ARGUMENT, "k=%u outside [1, 64]", k);
    if (k + cr_hamming_parity_bits(k) + 1 > 121) return fail(CR_ERR_UNSUPPORTED, "codeword too long");
    const uint32_t N = k + cr_hamming_parity_bits(k) + 1;  // codeword bits incl. overall parity
    const uint32_t cc = N + 1;
    std::vector<unsigned __int128> c(cc, 0);
    // hamming.cu:183-208
    c[0] = (unsigned __int128)1 << k;
    c[1] = with_1bit ? ((unsigned __int128)1 << k) * N : 0;
    c[2] = 0;
    for (uint32_t i = 4; i < cc; i += 2) {
        c[i] = (unsigned __int128)weights[i] + weights[i - 1];
        c[i - 1] = 0;
    }
    if (with_1bit)
        for (uint32_t i = 3; i < cc; i += 2)
            c[i] = i + 1 < cc ? (unsigned __int128)(i + 1) * c[i + 1] + (unsigned __int128)(N - i + 1) * c[i - 1]
                              : (unsigned __int128)(N - i + 1) * c[i - 1];
    for (uint32_t i = 3; i < cc; ++i) {
        if (c[i] >> (128 - k))
            return fail(CR_ERR_UNSUPPORTED, "reported count of weight %u exceeds 128 bits at k=%u", i, k);
        c[i] <<= k;
    }
    for (uint32_t i = 0; i < cc; ++i) {
        out_u128[2 * i] = (uint64_t)c[i];
        out_u128[2 * i + 1] = (uint64_t)(c[i] >> 64);
    }
    return CR_OK;
}

namespace {

// Histograms of the sample ranges [bounds[i], bounds[i+1]), i < count, in one
// launch: hists[i * nbins + b].  Empty ranges stay zero.
cr_status an_mc_ranges(cr_context *ctx, uint64_t A, uint32_t n, const uint64_t *bounds, uint32_t count,
                       uint64_t seed, uint64_t *hists, uint32_t nbins, double *kernel_ms) {
    if (!ctx || !hists || !bounds) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (A == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A must be >= 1");
    if (n < 1 || n > 63) return fail(CR_ERR_UNSUPPORTED, "AN Monte Carlo supports 1 <= n <= 63 (got %u)", n);
    if ((unsigned __int128)A * ((1ull << n) - 1) >> 64)
        return fail(CR_ERR_UNSUPPORTED, "codewords A*(2^n-1) exceed 64 bits");
    if (count < 1 || count > kMcMaxRanges)
        return fail(CR_ERR_INVALID_ARGUMENT, "%u sample ranges outside [1, %u]", count, kMcMaxRanges);
    for (uint32_t i = 0; i < count; ++i)
        if (bounds[i + 1] < bounds[i]) return fail(CR_ERR_INVALID_ARGUMENT, "sample bounds decrease at %u", i + 1);
    uint32_t bits = 0;
    for (unsigned __int128 x = (unsigned __int128)A * ((1ull << n) - 1); x; x >>= 1) ++bits;
    const uint32_t need = bits + 1;
    if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but distances reach %u", nbins, need - 1);
    CR_HIP(hipSetDevice(ctx->device));
    memset(hists, 0, (size_t)count * nbins * sizeof(uint64_t));
    if (kernel_ms) *kernel_ms = 0;
    const uint64_t entries = (uint64_t)count * need;
    CR_TRY(ctx->hist.reserve(entries * sizeof(uint64_t)));
    CR_TRY(ctx->counters.reserve(sizeof(uint64_t)));
    CR_TRY(ctx->jobs.reserve(count * sizeof(McSegment)));
    unsigned long long *out = ctx->hist.as<unsigned long long>();
    std::vector<McSegment> segs;
    for (uint32_t i = 0; i < count; ++i)
        if (bounds[i + 1] > bounds[i]) segs.push_back({bounds[i], bounds[i + 1], 0, 0, out + (uint64_t)i * need});
    if (segs.empty()) return CR_OK;
    CR_TRY(fill_zero_u64(out, entries, ctx->stream));
    CR_TRY(launch_an_mc(A, n, seed, segs.data(), (uint32_t)segs.size(), ctx->jobs.as<McSegment>(),
                        ctx->counters.as<unsigned long long>(), need, ctx->cus, ctx->stream,
                        kernel_ms ? ctx->ev0 : nullptr, kernel_ms ? ctx->ev1 : nullptr));
    CR_TRY(ctx->host.reserve(entries * sizeof(uint64_t)));
    CR_HIP(hipMemcpyAsync(ctx->host.ptr, out, entries * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t *h = ctx->host.as<uint64_t>();
    for (uint32_t i = 0; i < count; ++i) memcpy(hists + (uint64_t)i * nbins, h + (uint64_t)i * need, need * sizeof(uint64_t));
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}

}  // namespace

extern "C" cr_status cr_an_mc(cr_context *ctx, uint64_t A, uint32_t n, uint64_t m, uint64_t seed, uint64_t *hist,
                              uint32_t nbins, double *kernel_ms) {
    clear_error();
    const uint64_t bounds[2] = {0, m};
    return an_mc_ranges(ctx, A, n, bounds, 1, seed, hist, nbins, kernel_ms);
}

extern "C" cr_status cr_an_mc_ranges(cr_context *ctx, uint64_t A, uint32_t n, const uint64_t *bounds,
                                     uint32_t count, uint64_t seed, uint64_t *hists, uint32_t nbins,
                                     double *kernel_ms) {
    clear_error();
    return an_mc_ranges(ctx, A, n, bounds, count, seed, hists, nbins, kernel_ms);
}
