// Context management, error reporting and the exhaustive-AN entry point.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {

static thread_local char g_err[1024];

cr_status fail(cr_status s, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return s;
}

void clear_error() { g_err[0] = 0; }

cr_status DeviceBuffer::reserve(size_t want) {
    if (want <= bytes) return CR_OK;
    release();
    size_t n = std::max(want, (size_t)256);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess)
        return fail(CR_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    ptr = p;
    bytes = n;
    return CR_OK;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

cr_status HostBuffer::reserve(size_t want) {
    if (want <= bytes) return CR_OK;
    release();
    size_t n = std::max(want, (size_t)4096);
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(CR_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    ptr = p;
    bytes = n;
    return CR_OK;
}

void HostBuffer::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void auto_pair_tuning(uint64_t N, uint32_t resident_waves, uint32_t &R, uint32_t &Q) {
    // rows per lane: small codes favour more, smaller row blocks (measured on
    // single-code launches: R = 2 up to 2^14 words, 4 at 2^15, 8 from 2^16)
    if (R == 0) R = N >= (1ull << 16) ? 8 : (N >= (1ull << 15) ? 4 : 2);
    if (Q == 0) {
        const uint64_t RB = 64ull * R;
        const uint64_t T = (N + RB - 1) / RB;
        const uint64_t target = 16ull * std::max<uint32_t>(resident_waves, 1);
        const long double q = (long double)T * (long double)T / (2.0L * (long double)target);
        uint64_t qq = q < 1 ? 1 : (uint64_t)q;
        // a lane adds at most R per column: keep R*CW well below 2^32
        const uint64_t qmax = ((1ull << 30) / R) / RB;
        Q = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(qq, 1), std::max<uint64_t>(qmax, 1));
    }
}



}  // namespace cr

using namespace cr;

extern "C" {

const char *cr_status_string(cr_status s) {
    switch (s) {
    case CR_OK: return "ok";
    case CR_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CR_ERR_UNSUPPORTED: return "unsupported";
    case CR_ERR_NO_DEVICE: return "no device";
    case CR_ERR_OUT_OF_MEMORY: return "out of memory";
    case CR_ERR_DEVICE: return "device error";
    case CR_ERR_ADMISSION: return "admission refused";
    case CR_ERR_CANCELLED: return "cancelled";
    case CR_ERR_NOT_FOUND: return "not found";
    case CR_ERR_BUSY: return "busy";
    }
    return "unknown status";
}

const char *cr_last_error(void) { return g_err; }

const char *cr_version(void) { return "coding_reliability_amd 0.1 (gfx950)"; }

cr_status cr_context_create(int device, cr_context **out) {
    clear_error();
    if (!out) return fail(CR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(CR_ERR_NO_DEVICE, "device %d not in [0, %d)", device, n);
    hipDeviceProp_t prop;
    CR_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(CR_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    CR_HIP(hipSetDevice(device));
    // the device's copy of the allocation-free entry point's scratch array; a
    // failure here is reported by that entry point, which retries the look-up
    if (noalloc_setup_device(device) != CR_OK) clear_error();
    cr_context *c = new cr_context();
    c->device = device;
    c->cus = prop.multiProcessorCount;
    c->clock_mhz = prop.clockRate / 1000;
    c->hbm = prop.totalGlobalMem;
    snprintf(c->name, sizeof(c->name), "%s", prop.name);
    snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        cr_context_destroy(c);
        return fail(CR_ERR_DEVICE, "stream/event creation failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return CR_OK;
}

cr_status cr_context_destroy(cr_context *ctx) {
    if (!ctx) return CR_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->jobs.release();
    ctx->counters.release();
    ctx->hist.release();
    ctx->table_a.release();
    ctx->table_b.release();
    ctx->host.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CR_OK;
}

cr_status cr_device_memory(cr_context *ctx, uint64_t *free_bytes, uint64_t *total_bytes) {
    clear_error();
    if (!ctx || !free_bytes || !total_bytes) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    CR_HIP(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    CR_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return CR_OK;
}

cr_status cr_device_query(cr_context *ctx, cr_device_info *out) {
    clear_error();
    if (!ctx || !out) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", ctx->name);
    snprintf(out->arch, sizeof(out->arch), "%s", ctx->arch);
    out->compute_units = ctx->cus;
    out->max_clock_mhz = ctx->clock_mhz;
    out->hbm_bytes = ctx->hbm;
    return CR_OK;
}

cr_status cr_an_pairs_tuned(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                            const cr_pair_tuning *tuning, double *kernel_ms) {
    clear_error();
    if (!ctx || !raw) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (A == 0 || num_words == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A and num_words must be >= 1");
    if ((unsigned __int128)(num_words - 1) * A >> 64)
        return fail(CR_ERR_UNSUPPORTED, "codewords (num_words-1)*A exceed 64 bits");
    const uint32_t need = pair_max_distance(A, num_words) + 1;
    if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but distances reach %u", nbins, need - 1);
    CR_HIP(hipSetDevice(ctx->device));

    uint32_t R = tuning ? tuning->rows_per_lane : 0, Q = tuning ? tuning->chunk_blocks : 0;
    PairLaunchConfig cfg;
    cfg.lds_rows = need + 1;
    uint32_t Rr = R ? R : 8;
    const uint32_t per_cu = pair_waves_per_cu(Rr, cfg.lds_rows);
    auto_pair_tuning(num_words, per_cu * (uint32_t)ctx->cus, R, Q);
    cfg.R = R;
    cfg.waves = tuning ? tuning->waves : 0;
    // below about 100 us of work (1e9 pairs at 9.6e12 pairs/s) the queue's
    // serialized atomics, one per item, cost more than a static assignment's
    // imbalance
    const long double pairs = (long double)num_words * (long double)(num_words - 1) / 2;
    cfg.allow_static = cfg.waves == 0 && pairs < kStaticPairs;

    PairJob job{};
    job.A = A;
    job.N = num_words;
    job.mode = pair_mode(A, num_words);
    job.nbins = need;
    pair_job_geometry(job, R, Q);

    // device: the histogram (as kHistCopies line-aligned copies when many
    // short-lived waves flush it) then the work-queue head, cleared by one
    // fill; host (pinned): the job descriptor at 0, the copies from 256
    constexpr size_t kHostResult = 256;
    static_assert(sizeof(PairJob) <= kHostResult, "job staging");
    job.out_copies = cfg.allow_static ? dev::kHistCopies : 1;
    const uint64_t cstride = dev::hist_copy_stride(need);
    const uint64_t entries = job.out_copies > 1 ? cstride * job.out_copies : need;
    CR_TRY(ctx->hist.reserve((entries + 1) * sizeof(uint64_t)));
    CR_TRY(ctx->jobs.reserve(sizeof(PairJob)));
    CR_TRY(ctx->host.reserve(kHostResult + entries * sizeof(uint64_t)));
    job.out = ctx->hist.as<unsigned long long>();
    unsigned long long *queue = job.out + entries;
    job.item_begin = 0;
    memcpy(ctx->host.ptr, &job, sizeof(job));
    CR_TRY(fill_zero_u64(job.out, entries + 1, ctx->stream));
    CR_HIP(hipMemcpyAsync(ctx->jobs.ptr, ctx->host.ptr, sizeof(PairJob), hipMemcpyHostToDevice, ctx->stream));
    cfg.zero_counter = false;
    CR_TRY(launch_pairs(ctx->jobs.as<PairJob>(), 1, job.items, queue, cfg, ctx->cus, ctx->stream,
                        kernel_ms ? ctx->ev0 : nullptr, kernel_ms ? ctx->ev1 : nullptr));
    const uint64_t *res = reinterpret_cast<const uint64_t *>(ctx->host.as<char>() + kHostResult);
    CR_HIP(hipMemcpyAsync((void *)res, job.out, entries * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    memset(raw, 0, nbins * sizeof(uint64_t));
    for (uint32_t b = 0; b < need; ++b) {
        uint64_t s = 0;
        for (uint32_t c = 0; c < job.out_copies; ++c) s += res[c * cstride + b];
        raw[b] = s;
    }
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}

cr_status cr_an_pairs(cr_context *ctx, uint64_t A, uint64_t num_words, uint64_t *raw, uint32_t nbins,
                      double *kernel_ms) {
    // the same histogram from the code's spectrum when that is clearly cheaper
    uint32_t k = 0;
    if (spectral_preferred(A, num_words, &k)) return cr_an_pairs_spectral(ctx, A, k, raw, nbins, kernel_ms);
    return cr_an_pairs_tuned(ctx, A, num_words, raw, nbins, nullptr, kernel_ms);
}

}  // extern "C"
