// SDC table of the (N,k) Extended Hamming codes in the layout of the legacy
// table program (legacy/hamming_count_undetectable_errors.cpp:226-339).
//
// The table is derived from the weight distribution W[w] of the non-extended
// Hamming code (data bits plus the r parity bits of the non-zero masks):
//   ext[0] = W[0], ext[odd] = 0, ext[even i] = W[i] + W[i-1]      (:233-241)
//   cor[0] = ext[0], cor[even] = ext, cor[odd i] = ext[i-1]*(N-(i-1)) + ext[i+1]*(i+1)   (:258-266)
//   totals = x 2^k, maximum transitions 2^k * C(N, i), probabilities = totals / maximum.
// W comes from the device (exhaustive weight histogram, feasible up to k = 40)
// or, for any k <= 64, from the MacWilliams identity: the dual of the
// non-extended code is spanned by the r rows [mask_t | e_t], so
//   W[w] = 2^-r * sum over the 2^r dual words u of K_w(wt(u)),
//   K_w(x) = sum_j (-1)^j C(x, j) C(n - x, w - j),   n = k + r,
// evaluated exactly in 128-bit integers (|terms| < 2^75 for n <= 71).
// Totals reach 2^133 for (72,64); they are kept in 256-bit integers as the
// legacy program does (boost uint256_t).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "cr_internal.h"

using namespace cr;

namespace {

struct U256 {
    uint64_t w[4] = {0, 0, 0, 0};
};

U256 u256(unsigned __int128 v) {
    U256 r;
    r.w[0] = (uint64_t)v;
    r.w[1] = (uint64_t)(v >> 64);
    return r;
}

U256 add(const U256 &a, const U256 &b) {
    U256 r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
        c += (unsigned __int128)a.w[i] + b.w[i];
        r.w[i] = (uint64_t)c;
        c >>= 64;
    }
    return r;
}

U256 mul_small(const U256 &a, uint64_t m) {
    U256 r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
        c += (unsigned __int128)a.w[i] * m;
        r.w[i] = (uint64_t)c;
        c >>= 64;
    }
    return r;
}

U256 shl(const U256 &a, unsigned s) {
    U256 r;
    const unsigned q = s / 64, b = s % 64;
    for (int i = 3; i >= 0; --i) {
        const int src = i - (int)q;
        if (src < 0) continue;
        uint64_t v = a.w[src] << b;
        if (b && src > 0) v |= a.w[src - 1] >> (64 - b);
        r.w[i] = v;
    }
    return r;
}

bool less(const U256 &a, const U256 &b) {
    for (int i = 3; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return false;
}

bool is_zero(const U256 &a) { return !(a.w[0] | a.w[1] | a.w[2] | a.w[3]); }

// a / d (d < 2^64), remainder returned
uint64_t divmod(U256 &a, uint64_t d) {
    unsigned __int128 r = 0;
    for (int i = 3; i >= 0; --i) {
        r = (r << 64) | a.w[i];
        a.w[i] = (uint64_t)(r / d);
        r %= d;
    }
    return (uint64_t)r;
}

std::string dec(U256 a) {
    if (is_zero(a)) return "0";
    std::vector<uint64_t> parts;
    while (!is_zero(a)) parts.push_back(divmod(a, 10000000000000000000ull));
    std::string s = std::to_string(parts.back());
    char buf[32];
    for (int i = (int)parts.size() - 2; i >= 0; --i) {
        snprintf(buf, sizeof(buf), "%019llu", (unsigned long long)parts[i]);
        s += buf;
    }
    return s;
}

// the legacy width rule: number of decimal digits, at least one
size_t width(const U256 &a) { return dec(a).size(); }

double to_double(const U256 &a) {
    long double v = 0;
    for (int i = 3; i >= 0; --i) v = v * 18446744073709551616.0L + (long double)a.w[i];
    return (double)v;
}

U256 binom(uint32_t n, uint32_t k) {
    if (k > n) return U256{};
    if (k > n - k) k = n - k;
    unsigned __int128 b = 1;  // C(n, k) < 2^70 for n <= 73
    for (uint32_t i = 1; i <= k; ++i) b = b * (n - (k - i)) / i;
    return u256(b);
}

// Non-extended weight distribution by the MacWilliams identity (see header).
std::vector<unsigned __int128> weights_macwilliams(uint32_t k, const uint64_t masks[7], uint32_t r) {
    const uint32_t n = k + r;
    // binomials up to n in signed 128-bit
    std::vector<std::vector<__int128>> C(n + 1, std::vector<__int128>(n + 1, 0));
    for (uint32_t a = 0; a <= n; ++a) {
        C[a][0] = 1;
        for (uint32_t b = 1; b <= a; ++b) C[a][b] = C[a - 1][b - 1] + (b < a ? C[a - 1][b] : 0);
    }
    std::vector<__int128> acc(n + 1, 0);
    for (uint32_t T = 0; T < (1u << r); ++T) {
        uint64_t x = 0;
        for (uint32_t t = 0; t < r; ++t)
            if (T >> t & 1) x ^= masks[t];
        const uint32_t wt = (uint32_t)__builtin_popcountll(x) + (uint32_t)__builtin_popcount(T);
        for (uint32_t w = 0; w <= n; ++w) {
            __int128 kw = 0;
            for (uint32_t j = 0; j <= w && j <= wt; ++j) {
                if (w - j > n - wt) continue;
                const __int128 term = C[wt][j] * C[n - wt][w - j];
                kw += (j & 1) ? -term : term;
            }
            acc[w] += kw;
        }
    }
    std::vector<unsigned __int128> W(n + 1, 0);
    for (uint32_t w = 0; w <= n; ++w) W[w] = (unsigned __int128)(acc[w] >> r);
    return W;
}

// 1234567 ns -> "1.234.567" (the legacy StopWatch print)
std::string legacy_ns(uint64_t ns) {
    uint64_t max = 1000;
    while (ns / max > 0) max *= 1000;
    max /= 1000;
    std::ostringstream o;
    o << ns / max;
    char buf[8];
    while (max > 1) {
        ns %= max;
        max /= 1000;
        snprintf(buf, sizeof(buf), ".%03llu", (unsigned long long)(ns / max));
        o << buf;
    }
    return o.str();
}

}  // namespace

extern "C" cr_status cr_hamming_sdc_table(cr_context *ctx, uint32_t k, int method, uint64_t *weights_u128,
                                          uint32_t nweights, char **text) {
    clear_error();
    if (!text) return fail(CR_ERR_INVALID_ARGUMENT, "NULL text");
    *text = nullptr;
    if (k < 1 || k > 64) return fail(CR_ERR_INVALID_ARGUMENT, "k=%u outside [1, 64]", k);
    if (method < 0 || method > 2) return fail(CR_ERR_INVALID_ARGUMENT, "method %d", method);
    if (method == 0) method = k <= 40 ? 1 : 2;
    if (method == 1 && k > 40) return fail(CR_ERR_UNSUPPORTED, "device enumeration of 2^%u words", k);
    if (method == 1 && !ctx) return fail(CR_ERR_INVALID_ARGUMENT, "NULL context");
    uint64_t masks[7];
    const uint32_t r = hamming_masks(k, masks);
    const uint32_t N = k + r + 1, counts = N + 1;
    if (weights_u128 && nweights < counts)
        return fail(CR_ERR_INVALID_ARGUMENT, "nweights=%u < %u", nweights, counts);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<U256> raw(counts);
    if (method == 1) {
        cr_ham_request q{};
        q.k = k;
        q.source = CR_HAM_RANGE;
        q.first = 0;
        q.count = 1ull << k;
        std::vector<uint64_t> h(counts, 0);
        CR_TRY(cr_hamming_weights(ctx, &q, h.data(), counts, nullptr));
        for (uint32_t i = 0; i < counts; ++i) raw[i] = u256(h[i]);
    } else {
        const auto W = weights_macwilliams(k, masks, r);
        for (uint32_t i = 0; i < W.size() && i < counts; ++i) raw[i] = u256(W[i]);
    }
    const uint64_t ns =
        (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    if (weights_u128)
        for (uint32_t i = 0; i < counts; ++i) {
            weights_u128[2 * i] = raw[i].w[0];
            weights_u128[2 * i + 1] = raw[i].w[1];
        }

    std::vector<U256> ext(counts), cor(counts), ext_t(counts), cor_t(counts);
    U256 max_raw, max_ext, max_cor, max_ext_t, max_cor_t;
    ext[0] = raw[0];
    for (uint32_t i = 1; i < counts; ++i) {
        ext[i] = (i & 1) ? U256{} : add(raw[i], raw[i - 1]);
        if (less(max_raw, raw[i])) max_raw = raw[i];
        if (less(max_ext, ext[i])) max_ext = ext[i];
    }
    cor[0] = ext[0];
    for (uint32_t i = 1; i < counts; ++i) {
        if (i & 1) {
            cor[i] = mul_small(ext[i - 1], N - (i - 1));
            if (i + 1 < counts) cor[i] = add(cor[i], mul_small(ext[i + 1], i + 1));
        } else {
            cor[i] = ext[i];
        }
        if (less(max_cor, cor[i])) max_cor = cor[i];
    }
    for (uint32_t i = 0; i < counts; ++i) {
        ext_t[i] = shl(ext[i], k);
        cor_t[i] = shl(cor[i], k);
        if (less(max_ext_t, ext_t[i])) max_ext_t = ext_t[i];
        if (less(max_cor_t, cor_t[i])) max_cor_t = cor_t[i];
    }
    const size_t w_raw = width(max_raw), w_ext = width(max_ext), w_cor = width(max_cor), w_ext_t = width(max_ext_t),
                 w_cor_t = width(max_cor_t), w_tra = width(shl(binom(N, N / 2), k));

    const bool pow2 = (N & (~N + 1)) == N;  // the legacy "shortened" label test
    const char *sh = pow2 ? "shortened " : "";
    std::ostringstream o;
    o << "# SDC probabilities for a (n,k) = (" << N << ',' << k << ") " << sh << "correcting Extended Hamming code\n";
    o << "# The reported colunms are:\n";
    o << "#   1) bit width\n";
    o << "#   2) weight distribution of " << sh << "detecting-only Hamming\n";
    o << "#   3) weight distribution of " << sh << "detecting-only Extended Hamming\n";
    o << "#   4) weight distribution of " << sh << "correcting Extended Hamming\n";
    o << "#   5) error pattern weight distribution of " << sh << "detecting-only Extended Hamming\n";
    o << "#   6) error pattern weight distribution of " << sh << "correcting Extended Hamming\n";
    o << "#   7) maximum # transitions of " << sh << "extended Hamming (2^k * (n over col 1))\n";
    o << "#   8) probability of " << sh << "detecting-only Extended Hamming (col 5 / col 7)\n";
    o << "#   9) probability of " << sh << "correcting Extended Hamming (col 6 / col 7)\n";
    o << "#Distances:\n";
    auto pad = [](const std::string &s, size_t w) { return s.size() >= w ? s : std::string(w - s.size(), ' ') + s; };
    U256 total, transitions;
    char prob[64];
    for (uint32_t i = 0; i < counts; ++i) {
        const U256 maxt = shl(binom(N, i), k);
        total = add(total, raw[i]);
        transitions = add(transitions, cor_t[i]);
        o << pad(std::to_string(i), 4) << ',' << pad(dec(raw[i]), w_raw) << ',' << pad(dec(ext[i]), w_ext) << ','
          << pad(dec(cor[i]), w_cor) << ',' << pad(dec(ext_t[i]), w_ext_t) << ',' << pad(dec(cor_t[i]), w_cor_t) << ','
          << pad(dec(maxt), w_tra);
        snprintf(prob, sizeof(prob), ",%.10f", to_double(ext_t[i]) / to_double(maxt));
        o << prob;
        snprintf(prob, sizeof(prob), ",%.10f\n", to_double(cor_t[i]) / to_double(maxt));
        o << prob;
    }
    if (dec(total) != dec(shl(u256(1), k)))
        return fail(CR_ERR_DEVICE, "(%u/%u): weight total %s != 2^%u", N, k, dec(total).c_str(), k);
    o << "# (" << N << '/' << k << ") : Computation took " << legacy_ns(ns) << "ns for " << dec(shl(u256(1), k))
      << " code words and " << dec(transitions) << " transitions.\n\n\n";
    const std::string s = o.str();
    char *out = static_cast<char *>(malloc(s.size() + 1));
    if (!out) return fail(CR_ERR_OUT_OF_MEMORY, "table text");
    memcpy(out, s.c_str(), s.size() + 1);
    *text = out;
    return CR_OK;
}
