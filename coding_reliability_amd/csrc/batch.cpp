// Batched exhaustive AN histograms: many codes, one persistent launch,
// optionally replayed from a captured hipGraph; plus the autotuner.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "cr_internal.h"

namespace cr {

namespace {

struct BatchLayout {
    std::vector<PairJob> jobs;  // longest code first
    PairLaunchConfig cfg;
    uint64_t total_items = 0;
};

// Builds the work list: per code a PairJob writing to d_hist + c*stride, the
// chunk width chosen so that items carry about total_pairs/(items_per_wave *
// resident waves) pairs, jobs sorted by pair count (longest first).
cr_status build_batch(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint32_t stride,
                      const cr_pair_tuning *tuning, unsigned long long *d_hist, BatchLayout &L) {
    if (!codes || ncodes == 0) return fail(CR_ERR_INVALID_ARGUMENT, "empty batch");
    uint32_t R = tuning && tuning->rows_per_lane ? tuning->rows_per_lane : 8;
    const uint32_t ipw = tuning && tuning->items_per_wave ? tuning->items_per_wave : 16;
    uint32_t max_need = 0;
    long double total_pairs = 0;
    for (uint32_t c = 0; c < ncodes; ++c) {
        const uint64_t A = codes[c].A, N = codes[c].num_words;
        if (A == 0 || N == 0) return fail(CR_ERR_INVALID_ARGUMENT, "code %u: A and num_words must be >= 1", c);
        if ((unsigned __int128)(N - 1) * A >> 64) return fail(CR_ERR_UNSUPPORTED, "code %u exceeds 64-bit codewords", c);
        const uint32_t need = pair_max_distance(A, N) + 1;
        if (need > stride) return fail(CR_ERR_INVALID_ARGUMENT, "code %u needs %u bins > stride %u", c, need, stride);
        max_need = std::max(max_need, need);
        total_pairs += (long double)N * (long double)(N - 1) / 2;
    }
    L.cfg.R = R;
    L.cfg.lds_rows = max_need + 1;
    L.cfg.waves = tuning ? tuning->waves : 0;
    const uint32_t per_cu = pair_waves_per_cu(R, L.cfg.lds_rows);
    if (per_cu == 0) return fail(CR_ERR_INVALID_ARGUMENT, "rows_per_lane must be 2, 4, 8 or 16 (got %u)", R);
    const long double waves = (long double)per_cu * ctx->cus;
    const long double target = std::max<long double>(total_pairs / (ipw * waves), 1.0L);
    const uint64_t RB = 64ull * R;
    const uint64_t qmax = std::max<uint64_t>(((1ull << 30) / R) / RB, 1);
    L.jobs.resize(ncodes);
    for (uint32_t c = 0; c < ncodes; ++c) {
        PairJob &j = L.jobs[c];
        j = PairJob{};
        j.A = codes[c].A;
        j.N = codes[c].num_words;
        j.mode = pair_mode(j.A, j.N);
        j.nbins = pair_max_distance(j.A, j.N) + 1;
        j.out = d_hist + (size_t)c * stride;
        uint64_t Q = (uint64_t)(target / ((long double)RB * RB));
        Q = std::min<uint64_t>(std::max<uint64_t>(Q, 1), qmax);
        pair_job_geometry(j, R, (uint32_t)Q);
    }
    std::stable_sort(L.jobs.begin(), L.jobs.end(), [](const PairJob &a, const PairJob &b) {
        return (unsigned __int128)a.N * a.N > (unsigned __int128)b.N * b.N;
    });
    uint64_t total = 0;
    for (auto &j : L.jobs) {
        j.item_begin = total;
        total += j.items;
    }
    L.total_items = total;
    return CR_OK;
}

}  // namespace

}  // namespace cr

using namespace cr;

struct cr_batch_plan {
    cr_context *ctx = nullptr;
    uint32_t ncodes = 0, stride = 0;
    DeviceBuffer jobs, hist, counter;
    HostBuffer host;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

cr_status cr_an_batch(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint64_t *raw, uint32_t stride,
                      const cr_pair_tuning *tuning, double *kernel_ms) {
    clear_error();
    if (!ctx || !raw) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    CR_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)ncodes * stride;
    CR_TRY(ctx->hist.reserve(n * sizeof(uint64_t)));
    BatchLayout L;
    CR_TRY(build_batch(ctx, codes, ncodes, stride, tuning, ctx->hist.as<unsigned long long>(), L));
    CR_TRY(ctx->jobs.reserve(L.jobs.size() * sizeof(PairJob)));
    CR_TRY(ctx->counters.reserve(sizeof(uint64_t)));
    CR_TRY(fill_zero_u64(ctx->hist.as<unsigned long long>(), n, ctx->stream));
    CR_HIP(hipMemcpyAsync(ctx->jobs.ptr, L.jobs.data(), L.jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice,
                          ctx->stream));
    CR_TRY(launch_pairs(ctx->jobs.as<PairJob>(), (uint32_t)L.jobs.size(), L.total_items,
                        ctx->counters.as<unsigned long long>(), L.cfg, ctx->cus, ctx->stream,
                        kernel_ms ? ctx->ev0 : nullptr, kernel_ms ? ctx->ev1 : nullptr));
    CR_TRY(ctx->host.reserve(n * sizeof(uint64_t)));
    CR_HIP(hipMemcpyAsync(ctx->host.ptr, ctx->hist.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(raw, ctx->host.ptr, n * sizeof(uint64_t));
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}

cr_status cr_batch_plan_destroy(cr_batch_plan *p) {
    if (!p) return CR_OK;
    if (p->ctx) (void)hipSetDevice(p->ctx->device);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->jobs.release();
    p->hist.release();
    p->counter.release();
    p->host.release();
    delete p;
    return CR_OK;
}

cr_status cr_batch_plan_create(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint32_t stride,
                               const cr_pair_tuning *tuning, cr_batch_plan **out) {
    clear_error();
    if (!ctx || !out) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    CR_HIP(hipSetDevice(ctx->device));
    cr_batch_plan *p = new cr_batch_plan();
    p->ctx = ctx;
    p->ncodes = ncodes;
    p->stride = stride;
    auto bail = [&](cr_status s) {
        cr_batch_plan_destroy(p);
        return s;
    };
    const size_t n = (size_t)ncodes * stride;
    cr_status s;
    if ((s = p->hist.reserve(n * sizeof(uint64_t))) != CR_OK) return bail(s);
    if ((s = p->counter.reserve(sizeof(uint64_t))) != CR_OK) return bail(s);
    if ((s = p->host.reserve(n * sizeof(uint64_t))) != CR_OK) return bail(s);
    BatchLayout L;
    if ((s = build_batch(ctx, codes, ncodes, stride, tuning, p->hist.as<unsigned long long>(), L)) != CR_OK)
        return bail(s);
    if ((s = p->jobs.reserve(L.jobs.size() * sizeof(PairJob))) != CR_OK) return bail(s);
    hipError_t e = hipMemcpy(p->jobs.ptr, L.jobs.data(), L.jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(fail(CR_ERR_DEVICE, "job upload failed: %s", hipGetErrorString(e)));
    // capture: zero histograms, zero queue head + pair kernel, copy back
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bail(fail(CR_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e)));
    cr_status cs = fill_zero_u64(p->hist.as<unsigned long long>(), n, ctx->stream);
    if (cs == CR_OK)
        cs = launch_pairs(p->jobs.as<PairJob>(), (uint32_t)L.jobs.size(), L.total_items,
                          p->counter.as<unsigned long long>(), L.cfg, ctx->cus, ctx->stream, nullptr, nullptr);
    if (cs == CR_OK) {
        e = hipMemcpyAsync(p->host.ptr, p->hist.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) cs = fail(CR_ERR_DEVICE, "capture of copy-back failed: %s", hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    p->graph = g;
    if (cs != CR_OK) return bail(cs);
    if (e != hipSuccess) return bail(fail(CR_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e)));
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(fail(CR_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e)));
    *out = p;
    return CR_OK;
}

cr_status cr_batch_plan_run(cr_batch_plan *p, uint64_t *raw, double *elapsed_ms) {
    clear_error();
    if (!p) return fail(CR_ERR_INVALID_ARGUMENT, "NULL plan");
    cr_context *ctx = p->ctx;
    CR_HIP(hipSetDevice(ctx->device));
    CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    CR_HIP(hipGraphLaunch(p->exec, ctx->stream));
    CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    if (raw) memcpy(raw, p->host.ptr, (size_t)p->ncodes * p->stride * sizeof(uint64_t));
    if (elapsed_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *elapsed_ms = ms;
    }
    return CR_OK;
}

cr_status cr_autotune_batch(cr_context *ctx, const cr_an_code *codes, uint32_t ncodes, uint32_t repeats,
                            cr_pair_tuning *best, double *best_ms) {
    clear_error();
    if (!ctx || !best) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (repeats == 0) repeats = 1;
    uint32_t stride = 0;
    for (uint32_t c = 0; c < ncodes; ++c) {
        if (codes[c].A == 0 || codes[c].num_words == 0) return fail(CR_ERR_INVALID_ARGUMENT, "bad code %u", c);
        stride = std::max(stride, pair_max_distance(codes[c].A, codes[c].num_words) + 1);
    }
    const uint32_t Rs[] = {4, 8, 16};
    const uint32_t ipws[] = {4, 16, 64};
    double bms = 1e300;
    cr_pair_tuning bt{};
    for (uint32_t R : Rs)
        for (uint32_t ipw : ipws) {
            cr_pair_tuning t{R, 0, 0, ipw};
            cr_batch_plan *p = nullptr;
            CR_TRY(cr_batch_plan_create(ctx, codes, ncodes, stride, &t, &p));
            double m = 1e300;
            for (uint32_t r = 0; r <= repeats; ++r) {  // first run warms up
                double ms = 0;
                cr_status s = cr_batch_plan_run(p, nullptr, &ms);
                if (s != CR_OK) {
                    cr_batch_plan_destroy(p);
                    return s;
                }
                if (r) m = std::min(m, ms);
            }
            cr_batch_plan_destroy(p);
            if (m < bms) {
                bms = m;
                bt = t;
            }
        }
    *best = bt;
    if (best_ms) *best_ms = bms;
    return CR_OK;
}

}  // extern "C"
