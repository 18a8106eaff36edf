// Host side of the spectral exhaustive AN histogram (spectral.hip): batching
// of the 2^(n-k) transforms, the exact weight sums W(w) and the Krawtchouk
// transform D(d) = 2^-n sum_w K_d(w) W(w).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {

size_t spectral_tile_count(uint32_t k);
// Resident workgroups per CU of the last pass launch_spectral_batch runs for
// (layout, k), from the occupancy API (4 if the query fails).
uint32_t spectral_last_pass_per_cu(uint32_t layout, uint32_t k);
cr_status launch_spectral_batch(void *X, uint32_t layout, uint32_t k, uint64_t A, uint32_t Ainv, uint64_t tc,
                                uint64_t s0, uint32_t nchunks, unsigned long long *slots, uint32_t nslots,
                                uint32_t nbins, hipStream_t stream);

namespace {

using i128 = __int128;
using u128 = unsigned __int128;

uint32_t bit_length(uint64_t x) { return x ? 64 - (uint32_t)__builtin_clzll(x) : 0; }

// Inverse of an odd a modulo 2^32 (Newton: each step doubles the correct bits).
uint32_t inverse_mod_2_32(uint32_t a) {
    uint32_t x = a;  // correct to 3 bits for odd a
    for (int i = 0; i < 5; ++i) x *= 2u - a * x;
    return x;
}

int64_t binom(uint32_t n, uint32_t r) {
    if (r > n) return 0;
    int64_t b = 1;
    for (uint32_t i = 1; i <= r; ++i) b = b * (int64_t)(n - r + i) / (int64_t)i;  // exact at every step
    return b;
}

// K_d(w) for length n: sum_j (-1)^j C(w, j) C(n - w, d - j).
i128 krawtchouk(uint32_t n, uint32_t d, uint32_t w) {
    i128 s = 0;
    for (uint32_t j = 0; j <= d && j <= w; ++j) {
        const i128 t = (i128)binom(w, j) * binom(n - w, d - j);
        s += (j & 1) ? -t : t;
    }
    return s;
}

// Odd part of A, codeword bits n and support of the spectral path at k.
bool spectral_shape(uint64_t A, uint32_t k, uint64_t *Ao, uint32_t *n) {
    if (A == 0 || k < 14 || k > 32) return false;
    const uint64_t N = 1ull << k;
    if ((unsigned __int128)(N - 1) * A >> 64) return false;
    uint64_t a = A;
    while ((a & 1) == 0) a >>= 1;
    if (a >> 32) return false;
    const uint32_t bits = std::max<uint32_t>(bit_length(a * (N - 1)), k);
    // 2^(n-k) transforms of 2^k: beyond n = 2k that is more work than the
    // 2^(2k-1) pairs themselves
    if (bits > 48 || bits > 2 * k) return false;
    *Ao = a;
    *n = bits;
    return true;
}

}  // namespace

bool spectral_preferred(uint64_t A, uint64_t num_words, uint32_t *k) {
    if (num_words == 0 || (num_words & (num_words - 1))) return false;
    const uint32_t kk = (uint32_t)__builtin_ctzll(num_words);
    uint64_t Ao = 0;
    uint32_t n = 0;
    if (!spectral_shape(A, kk, &Ao, &n)) return false;
    SpectralPlan p;
    if (!spectral_plan(A, kk, 256, &p)) return false;
    const double spectral_s = std::ldexp(p.sec_per_element, (int)n) + 50e-6;
    const double pair_s = std::ldexp(1.0, 2 * (int)kk - 1) / 9.6e12;
    *k = kk;
    return spectral_s < 0.8 * pair_s;
}

bool spectral_plan(uint64_t A, uint32_t k, int cus, SpectralPlan *p) {
    uint64_t Ao = 0;
    uint32_t n = 0;
    if (!spectral_shape(A, k, &Ao, &n)) return false;
    const uint64_t N = 1ull << k;
    p->Ao = Ao;
    p->k = k;
    p->n = n;
    p->nw = n + 1;  // weights 0..n
    p->Ainv = inverse_mod_2_32((uint32_t)Ao) & (uint32_t)(N - 1);
    p->tc = (Ao * p->Ainv - 1) >> k;  // Ao * Ainv = 1 + tc * 2^k (below 2^56)
    p->chunks = 1ull << (n - k);
    // k = 17 .. 24 with at least two transforms (four at k = 17, whose last
    // pass needs 16 tiles of 2^15 elements per batch): pairs of transforms
    // through int16 (an even number per batch, starting at an even index)
    p->layout = (k >= 17 && k <= 24 && p->chunks >= (k == 17 ? 4u : 2u)) ? kSpectralInt16Pairs : kSpectralInt32;
    // measured seconds per transformed element: k = 17 .. 24 through int16
    // (0.80-0.84 ps: 2^40 elements in 0.89-0.92 s at k = 24, 2^37 in 109 ms
    // at k = 21, 2^34 in 13.8 ms at k = 18), two passes
    // through int32 up to k = 24 (2^40 in 2.35 s), three beyond (2^38 in
    // 0.99-1.02 s at k = 32)
    p->sec_per_element = p->layout == kSpectralInt16Pairs ? 0.84e-12 : (k <= 24 ? 2.2e-12 : 3.7e-12);
    // chunks of 2^k per launch: 256 MB of transforms in flight, about the
    // size of the last-level cache (measured at k = 24, A = 58659 through
    // int32: 2.36 s with 256 MB, 2.63 s with 128 MB, 2.67 s with 512 MB)
    const uint64_t chunk_bytes =
        N * ((k <= 16 || p->layout == kSpectralInt16Pairs) ? sizeof(int16_t) : sizeof(int32_t));
    p->per_batch = std::min<uint64_t>(p->chunks, std::max<uint64_t>(1, (1ull << 28) / chunk_bytes));
    p->batches = (p->chunks + p->per_batch - 1) / p->per_batch;
    p->buffer_bytes = p->per_batch * chunk_bytes;
    // one resident generation of last-pass workgroups (one slot each): at
    // k = 23 / 24 (two per CU) 4.7 % / 1.9 % faster than twice as many, at
    // k = 32 (int32) 3 % faster than half as many; a multiple of 16, as the
    // int16-pair pass needs
    p->nslots = ((uint32_t)cus * spectral_last_pass_per_cu(p->layout, k) + 15) & ~15u;
    p->slot_entries = (uint64_t)p->nslots * 4 * p->nw;  // per weight: lo sum, carries, hi sum, carries
    return true;
}

cr_status spectral_run(const SpectralPlan &p, void *X, unsigned long long *slots, uint64_t first_batch,
                       uint64_t nbatches, hipStream_t stream) {
    for (uint64_t b = first_batch; b < first_batch + nbatches && b < p.batches; ++b) {
        const uint64_t s0 = b * p.per_batch;
        const uint32_t nb = (uint32_t)std::min<uint64_t>(p.per_batch, p.chunks - s0);
        CR_TRY(launch_spectral_batch(X, p.layout, p.k, p.Ao, p.Ainv, p.tc, s0, nb, slots, p.nslots, p.nw, stream));
    }
    return CR_OK;
}

cr_status spectral_finish(const SpectralPlan &p, const uint64_t *sl, uint64_t *raw, uint32_t nbins) {
    // W(w) = sum over slots of lo + hi * 2^32, each a 64-bit sum plus a count
    // of carries out of it (exact in 128 bits: W <= 2^(n+k) <= 2^80)
    std::vector<u128> W(p.nw, 0);
    for (uint32_t q = 0; q < p.nslots; ++q)
        for (uint32_t w = 0; w < p.nw; ++w) {
            const uint64_t *e = sl + (uint64_t)q * 4 * p.nw + 4 * w;
            const u128 lo = (u128)e[0] + ((u128)e[1] << 64), hi = (u128)e[2] + ((u128)e[3] << 64);
            W[w] += lo + (hi << 32);
        }
    // the int16-pair passes sum F^2 / 4 (they transform the parities, not the signs)
    if (p.layout == kSpectralInt16Pairs)
        for (uint32_t w = 0; w < p.nw; ++w) W[w] *= 4;
    // D(d) = 2^-n sum_w K_d(w) W(w); partial sums stay below C(n, d) 2^(n+k) <= 2^125
    memset(raw, 0, nbins * sizeof(uint64_t));
    for (uint32_t d = 0; d < p.nw; ++d) {
        i128 s = 0;
        for (uint32_t w = 0; w < p.nw; ++w) s += krawtchouk(p.n, d, w) * (i128)W[w];
        const i128 unit = (i128)1 << p.n;
        if (s < 0 || s % unit != 0) return fail(CR_ERR_DEVICE, "spectral sums inconsistent at distance %u", d);
        const i128 D = s / unit;
        if (d == 0) {
            if (D != ((i128)1 << p.k)) return fail(CR_ERR_DEVICE, "spectral sums inconsistent at distance 0");
            continue;
        }
        if (D & 1) return fail(CR_ERR_DEVICE, "odd ordered-pair count at distance %u", d);
        raw[d] = (uint64_t)(D / 2);
    }
    return CR_OK;
}

}  // namespace cr

using namespace cr;

extern "C" cr_status cr_an_pairs_spectral(cr_context *ctx, uint64_t A, uint32_t k, uint64_t *raw, uint32_t nbins,
                                          double *kernel_ms) {
    clear_error();
    if (!ctx || !raw) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (A == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A must be >= 1");
    if (k < 14 || k > 32) return fail(CR_ERR_UNSUPPORTED, "spectral path: k=%u outside [14, 32]", k);
    const uint64_t N = 1ull << k;
    if ((unsigned __int128)(N - 1) * A >> 64) return fail(CR_ERR_UNSUPPORTED, "codewords (2^k-1)*A exceed 64 bits");
    const uint32_t need = pair_max_distance(A, N) + 1;
    if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but distances reach %u", nbins, need - 1);
    SpectralPlan p;
    if (!spectral_plan(A, k, ctx->cus, &p))
        return fail(CR_ERR_UNSUPPORTED, "spectral path: A=%llu at k=%u (odd part below 2^32, codewords of at most "
                    "min(2k, 48) bits)", (unsigned long long)A, k);
    CR_HIP(hipSetDevice(ctx->device));
    const cr_status st = [&]() -> cr_status {
        CR_TRY(ctx->table_a.reserve(p.buffer_bytes));
        CR_TRY(ctx->hist.reserve(p.slot_entries * sizeof(uint64_t)));
        CR_TRY(ctx->host.reserve(p.slot_entries * sizeof(uint64_t)));
        unsigned long long *slots = ctx->hist.as<unsigned long long>();
        CR_TRY(fill_zero_u64(slots, p.slot_entries, ctx->stream));
        if (kernel_ms) CR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        CR_TRY(spectral_run(p, ctx->table_a.ptr, slots, 0, p.batches, ctx->stream));
        if (kernel_ms) CR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        CR_HIP(hipMemcpyAsync(ctx->host.ptr, slots, p.slot_entries * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        CR_HIP(hipStreamSynchronize(ctx->stream));
        CR_TRY(spectral_finish(p, ctx->host.as<uint64_t>(), raw, nbins));
        if (kernel_ms) {
            float ms = 0;
            CR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            *kernel_ms = ms;
        }
        return CR_OK;
    }();
    // transform buffers above 1 GiB (16 GiB at k = 32) go back after the call:
    // a long-lived context (the reference entry points keep one per device)
    // would otherwise hold them until it is destroyed
    if (p.buffer_bytes > (1ull << 30)) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->table_a.release();
    }
    return st;
}
