// Reference-compatible grid approximation (run_ancoding_grid).
//
// The sample list of the reference kernels does not depend on the word, so it
// is produced once on the host with the same arithmetic as the reference
// kernel loops (IEEE double multiply, truncation to T, T = uint32/uint64
// wrap-around), uploaded, and consumed by the rectangle kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "cr_internal.h"

namespace cr {

namespace {

// Samples of dancoding_grid_1d_shared for one word
// (distance_distribution/src/an_coding_grid.cu:68,83-106): unrolled blocks of
// `unroll` samples while v < Aend_p, then single samples while v < Aend.
template <class T>
cr_status samples_1d(T A, uint32_t n, double step, T unroll, uint64_t cap, std::vector<uint64_t> &out) {
    const T Aend = (T)(((uint64_t)A) << n);
    const T Aend_p = Aend - A * (T)((double)unroll * step);
    T v = 0, k = 0;
    for (; v < Aend_p; k += unroll) {
        for (T u = 0; u < unroll; ++u, v = A * (T)((double)(T)(k + u) * step)) out.push_back(v);
        if (out.size() > cap) return fail(CR_ERR_INVALID_ARGUMENT, "grid sample list does not terminate");
    }
    for (; v < Aend; ++k, v = A * (T)((double)k * step)) {
        out.push_back(v);
        if (out.size() > cap) return fail(CR_ERR_INVALID_ARGUMENT, "grid sample list does not terminate");
    }
    return CR_OK;
}

// Samples of dancoding_grid_2d (an_coding_grid.cu:133): v = A*(T)(k*step) while v < Aend.
template <class T>
cr_status samples_2d(T A, uint32_t n, double step, uint64_t cap, std::vector<uint64_t> &out) {
    const T Aend = (T)(((uint64_t)A) << n);
    T v = 0, k = 0;
    for (; v < Aend; ++k, v = A * (T)((double)k * step)) {
        out.push_back(v);
        if (out.size() > cap) return fail(CR_ERR_INVALID_ARGUMENT, "grid sample list does not terminate");
    }
    return CR_OK;
}

uint32_t bitlen(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

struct GridCfg {
    uint64_t A, m1, m2, msgs;
    uint32_t n, count_counts;
};

// Host set-up of run_ancoding_grid (an_coding_grid.cu:217-261, one device).
cr_status grid_cfg(int gdim, uint32_t n, uint64_t m1, uint64_t m2, uint64_t A, GridCfg &g) {
    if (gdim != 1 && gdim != 2) return fail(CR_ERR_INVALID_ARGUMENT, "gdim must be 1 or 2");
    if (n < 1 || n > 48) return fail(CR_ERR_UNSUPPORTED, "n=%u outside [1, 48]", n);
    if (A == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A must be >= 1");
    if ((A & (A - 1)) == 0) A = 1;
    const uint64_t msgs = 1ull << n;
    if (A >= msgs) return fail(CR_ERR_INVALID_ARGUMENT, "A must be < 2^n (an_coding_grid.cu:219)");
    m1 = m1 > msgs ? msgs : m1;
    m2 = gdim == 1 ? msgs : (m2 > msgs ? msgs : m2);
    if (m1 == 0 || m2 == 0) return fail(CR_ERR_INVALID_ARGUMENT, "grid sizes must be >= 1");
    if (gdim == 1 && m1 < 3) return fail(CR_ERR_INVALID_ARGUMENT, "1D grid needs m >= 3 (an_coding_grid.cu:155)");
    const uint32_t h = (uint32_t)ceil(log((double)A) / log(2.0));
    g = GridCfg{A, m1, m2, msgs, n, n + h + 1};
    if ((unsigned __int128)A << n >> 64) return fail(CR_ERR_UNSUPPORTED, "A << n exceeds 64 bits");
    return CR_OK;
}

}  // namespace

}  // namespace cr

using namespace cr;

extern "C" cr_status cr_an_grid_ref(cr_context *ctx, int gdim, uint32_t n, uint64_t m1, uint64_t m2, uint64_t A,
                                    uint64_t *hist, uint32_t nbins, double *kernel_ms) {
    clear_error();
    if (!ctx || !hist) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    GridCfg g;
    CR_TRY(grid_cfg(gdim, n, m1, m2, A, g));
    CR_HIP(hipSetDevice(ctx->device));

    const double step = 1.0 * (double)g.msgs / (double)g.m1;
    const double step2 = 1.0 * (double)g.msgs / (double)g.m2;
    std::vector<uint64_t> cols, rows;
    const uint64_t cap = g.m1 + 64;
    bool t32;
    if (gdim == 1) {
        // T = uint32 iff count_counts <= 32 (an_coding_grid.cu:159-184)
        t32 = g.count_counts <= 32;
        const uint64_t unroll = g.m1 <= 16 ? 3 : 16;
        if (t32)
            CR_TRY(samples_1d<uint32_t>((uint32_t)g.A, n, step, (uint32_t)unroll, cap, cols));
        else
            CR_TRY(samples_1d<uint64_t>(g.A, n, step, unroll, cap, cols));
    } else {
        // T = uint32 iff Aend < 2^32 (an_coding_grid.cu:188-193)
        t32 = (g.A << n) < (1ull << 32);
        if (t32)
            CR_TRY(samples_2d<uint32_t>((uint32_t)g.A, n, step, cap, cols));
        else
            CR_TRY(samples_2d<uint64_t>(g.A, n, step, cap, cols));
        rows.resize(g.m2);
        for (uint64_t i = 0; i < g.m2; ++i) {
            if (t32)
                rows[i] = (uint32_t)((uint32_t)g.A * (uint32_t)((double)(uint32_t)i * step2));
            else
                rows[i] = g.A * (uint64_t)((double)i * step2);
        }
    }
    // highest possible distance: bit length of the OR of every word
    uint64_t orall = 0;
    for (uint64_t v : cols) orall |= v;
    if (gdim == 1) {
        // words A*i, i < 2^n: every bit up to the top bit of the largest word
        uint64_t mx = g.A * (g.msgs - 1);
        if (t32) mx = (uint32_t)mx;
        orall |= mx ? (~0ull >> __builtin_clzll(mx)) : 0;
    } else {
        for (uint64_t v : rows) orall |= v;
    }
    const uint32_t need = bitlen(orall) + 1;
    if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but distances reach %u", nbins, need - 1);

    CR_TRY(ctx->table_a.reserve(cols.size() * sizeof(uint64_t)));
    CR_HIP(hipMemcpyAsync(ctx->table_a.ptr, cols.data(), cols.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                          ctx->stream));
    if (!rows.empty()) {
        CR_TRY(ctx->table_b.reserve(rows.size() * sizeof(uint64_t)));
        CR_HIP(hipMemcpyAsync(ctx->table_b.ptr, rows.data(), rows.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    CR_TRY(ctx->hist.reserve(need * sizeof(uint64_t)));
    CR_TRY(ctx->counters.reserve(sizeof(uint64_t)));
    CR_TRY(fill_zero_u64(ctx->hist.as<unsigned long long>(), need, ctx->stream));

    RectLaunch L;
    L.A = g.A;
    L.row0 = 0;
    L.nrows = gdim == 1 ? g.msgs : g.m2;
    L.rows = gdim == 1 ? nullptr : ctx->table_b.as<uint64_t>();
    L.cols = ctx->table_a.as<uint64_t>();
    L.ncols = cols.size();
    L.mask32 = t32 ? 1 : 0;
    L.w2 = need - 1 > 32;
    L.counter = ctx->counters.as<unsigned long long>();
    L.out = ctx->hist.as<unsigned long long>();
    L.nbins = need;
    L.first_bin = gdim == 2 ? 1 : 0;  // dancoding_grid_2d flushes bins 1.. (an_coding_grid.cu:138)
    L.lds_rows = need + 1;
    CR_TRY(launch_rect(L, ctx->stream, kernel_ms ? ctx->ev0 : nullptr, kernel_ms ? ctx->ev1 : nullptr, ctx->cus));
    CR_TRY(ctx->host.reserve(need * sizeof(uint64_t)));
    CR_HIP(hipMemcpyAsync(ctx->host.ptr, L.out, need * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    CR_HIP(hipStreamSynchronize(ctx->stream));
    memset(hist, 0, nbins * sizeof(uint64_t));
    memcpy(hist, ctx->host.ptr, need * sizeof(uint64_t));
    if (kernel_ms) {
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *kernel_ms = ms;
    }
    return CR_OK;
}

extern "C" cr_status cr_an_grid_report(int gdim, uint32_t n, uint64_t m1, uint64_t m2, const uint64_t *hist,
                                       uint32_t count_counts, uint64_t *out_u128) {
    clear_error();
    if (!hist || !out_u128) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (gdim != 1 && gdim != 2) return fail(CR_ERR_INVALID_ARGUMENT, "gdim must be 1 or 2");
    if (n > 63) return fail(CR_ERR_UNSUPPORTED, "n=%u", n);
    const uint64_t msgs = 1ull << n;
    m1 = m1 > msgs ? msgs : m1;
    m2 = gdim == 1 ? msgs : (m2 > msgs ? msgs : m2);
    if (m1 == 0 || m2 == 0) return fail(CR_ERR_INVALID_ARGUMENT, "grid sizes must be >= 1");
    const long double factor =
        gdim == 1 ? powl(2.0L, (long double)n) / m1 : powl(4.0L, (long double)n) / (m1 * m2);
    for (uint32_t i = 0; i < count_counts; ++i) {
        const unsigned __int128 v = (unsigned __int128)(factor * hist[i]);
        out_u128[2 * i] = (uint64_t)v;
        out_u128[2 * i + 1] = (uint64_t)(v >> 64);
    }
    return CR_OK;
}
