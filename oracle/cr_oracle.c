/*
 * cr_oracle.c -- CPU oracle for the distance-distribution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the algorithms of
 * brics-db/coding_reliability that the HIP library implements, so that tests,
 * the smoke check and the CPU-baseline leg of bench.py can check the GPU
 * results.  Nothing in the product library links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Counts are kept in unsigned __int128
 * where the reference does (uint128_t, distance_distribution/src/globals.h:25).
 *
 * Pinning status of each function is documented in the header of oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

typedef unsigned __int128 u128;

#define OR_MAXBINS 130

static inline unsigned popc64(uint64_t x) { return (unsigned)__builtin_popcountll(x); }

/* ------------------------------------------------------------------------ */
/* Exhaustive AN coding                                                      */
/* ------------------------------------------------------------------------ */

/*
 * Restates countANCodingUndetectableErrors<64>
 * (distance_distribution/src/an_coding.cpp:50-102): 64-word row shards under
 * OpenMP schedule(dynamic,1); inside a shard the diagonal triangle adds 1 for
 * x==y and 2 for x<y (:68-75); the rectangle to the right adds 2 per pair
 * (:78-86); thread-local u64 bins merged with atomics (:95-98).
 *
 * Only row shards [shard_begin, shard_end) are processed, so a bounded sample
 * of a large problem can be timed (pass 0, ~0 for the full problem).  The
 * shard width is min(64, 2^n): the reference always walks 64 words, which for
 * n<6 reads words beyond the code (quirk not reproduced).
 * Returns the number of unordered pairs i<j evaluated.
 */
uint64_t or_an_cpu_counts(uint64_t n, uint64_t A, u128 *counts, uint64_t count_counts,
                          uint64_t shard_begin, uint64_t shard_end)
{
    const uint64_t msgs = 1ull << n;
    const uint64_t sz = msgs < 64 ? msgs : 64;
    const uint64_t nshards = msgs / sz;
    if (shard_end > nshards) shard_end = nshards;
    uint64_t pairs = 0;
#pragma omp parallel reduction(+ : pairs)
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(dynamic, 1)
        for (uint64_t s = shard_begin; s < shard_end; ++s) {
            const uint64_t shardX = s * sz;
            for (uint64_t x = 0; x < sz; ++x) {
                const uint64_t m1 = (shardX + x) * A;
                ++local[popc64(m1 ^ m1)];
                for (uint64_t y = x + 1; y < sz; ++y) {
                    const uint64_t m2 = (shardX + y) * A;
                    local[popc64(m1 ^ m2)] += 2;
                }
            }
            pairs += sz * (sz - 1) / 2;
            for (uint64_t shardY = shardX + sz; shardY < msgs; shardY += sz) {
                for (uint64_t x = 0; x < sz; ++x) {
                    const uint64_t m1 = (shardX + x) * A;
                    for (uint64_t y = 0; y < sz; ++y) {
                        const uint64_t m2 = (shardY + y) * A;
                        local[popc64(m1 ^ m2)] += 2;
                    }
                }
                pairs += sz * sz;
            }
        }
        for (uint64_t i = 0; i < count_counts && i < OR_MAXBINS; ++i) {
#pragma omp critical
            counts[i] += (u128)local[i];
        }
    }
    return pairs;
}

/*
 * Raw unordered-pair histogram raw[b] = #{0<=i<j<N : popcount(Ap*i ^ Ap*j) == b}
 * for an arbitrary word count N and 64-bit multiplier Ap (products taken mod
 * 2^64).  This is the quantity the reference GPU kernel dancoding_shared
 * accumulates (distance_distribution/src/an_coding.cu:68-97: v runs from
 * w+A up to A<<n, i.e. all j>i).  Plain brute force, row-parallel.
 */
void or_an_pairs_raw(uint64_t Ap, uint64_t N, uint64_t *raw, unsigned nbins)
{
    memset(raw, 0, nbins * sizeof(uint64_t));
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(dynamic, 16)
        for (uint64_t i = 0; i < N; ++i) {
            const uint64_t w = Ap * i;
            uint64_t v = w + Ap;
            for (uint64_t j = i + 1; j < N; ++j, v += Ap) ++local[popc64(w ^ v)];
        }
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            raw[b] += local[b];
        }
    }
}

/*
 * Post-processing of run_ancoding (distance_distribution/src/an_coding.cu:287-292):
 * out[0] = 2^n, out[i>=1] = raw[i] << 1 (in u128).
 */
void or_an_report(uint64_t n, const uint64_t *raw, unsigned count_counts, u128 *out)
{
    out[0] = (u128)1 << n;
    for (unsigned i = 1; i < count_counts; ++i) out[i] = ((u128)raw[i]) << 1;
}

/* minb/mincb search of run_ancoding (an_coding.cu:295-308), limit = count_counts/2;
 * the super-A driver uses (count_counts+1)/2 (super_a/src/main.cpp:157-170). */
void or_an_minb(const u128 *counts, unsigned limit, uint64_t *minb, u128 *mincb)
{
    *minb = 0xFFFF;
    *mincb = ~(u128)0;
    for (unsigned i = 1; i < limit; ++i)
        if (counts[i] != 0 && counts[i] < *mincb) {
            *minb = i;
            *mincb = counts[i];
            break;
        }
}

/* ------------------------------------------------------------------------ */
/* Reference grid (double step size)                                          */
/* ------------------------------------------------------------------------ */

/*
 * Sample list of dancoding_grid_1d_shared (distance_distribution/src/an_coding_grid.cu:60-107)
 * for one word: the loop visits v = A*(T)(k*step) for k = 0,1,2,...; the
 * unrolled body (Unroll samples, no bound check) runs while v < Aend_p with
 * Aend_p = Aend - A*(T)(Unroll*step) (:68,:83), then the remainder loop runs
 * while v < Aend (:102).  The list does not depend on the word, so it is
 * produced once.  T = uint32 (t32 != 0) or uint64 arithmetic, as in the
 * Caller dispatch (:159-184).  Returns the sample count (<= cap).
 */
uint64_t or_grid1d_ref_samples(uint64_t A, uint64_t n, double step, unsigned unroll, int t32,
                               uint64_t *out, uint64_t cap)
{
    uint64_t cnt = 0;
    if (t32) {
        const uint32_t A32 = (uint32_t)A, Aend = (uint32_t)(A << n);
        const uint32_t Aend_p = Aend - A32 * (uint32_t)((double)unroll * step);
        uint32_t v = 0, k = 0;
        for (; v < Aend_p; k += unroll) {
            for (uint32_t u = 0; u < unroll; ++u, v = A32 * (uint32_t)((double)(k + u) * step)) {
                if (cnt < cap) out[cnt] = v;
                ++cnt;
            }
        }
        for (; v < Aend; ++k, v = A32 * (uint32_t)((double)k * step)) {
            if (cnt < cap) out[cnt] = v;
            ++cnt;
        }
    } else {
        const uint64_t Aend = A << n;
        const uint64_t Aend_p = Aend - A * (uint64_t)((double)unroll * step);
        uint64_t v = 0, k = 0;
        for (; v < Aend_p; k += unroll) {
            for (uint64_t u = 0; u < unroll; ++u, v = A * (uint64_t)((double)(k + u) * step)) {
                if (cnt < cap) out[cnt] = v;
                ++cnt;
            }
        }
        for (; v < Aend; ++k, v = A * (uint64_t)((double)k * step)) {
            if (cnt < cap) out[cnt] = v;
            ++cnt;
        }
    }
    return cnt;
}

/*
 * Sample list of dancoding_grid_2d (an_coding_grid.cu:121-140):
 * v = A*(T)(k*step) for k = 0,1,... while v < Aend (no unrolling).
 */
uint64_t or_grid2d_ref_samples(uint64_t A, uint64_t n, double step, int t32, uint64_t *out, uint64_t cap)
{
    uint64_t cnt = 0;
    if (t32) {
        const uint32_t A32 = (uint32_t)A, Aend = (uint32_t)(A << n);
        uint32_t v = 0, k = 0;
        for (; v < Aend; ++k, v = A32 * (uint32_t)((double)k * step)) {
            if (cnt < cap) out[cnt] = v;
            ++cnt;
        }
    } else {
        const uint64_t Aend = A << n;
        uint64_t v = 0, k = 0;
        for (; v < Aend; ++k, v = A * (uint64_t)((double)k * step)) {
            if (cnt < cap) out[cnt] = v;
            ++cnt;
        }
    }
    return cnt;
}

/*
 * Rectangle histogram hist[b] += #{(r, s) : popcount(rows[r] ^ cols[s]) == b}
 * over explicit word lists (ordered pairs).  With rows = A*i for all i<2^n and
 * cols = the 1D sample list this is the kernel-level histogram of the
 * reference 1D grid; with rows = A*(T)(i*step2) it is the 2D grid
 * (an_coding_grid.cu:132).  If mask32, words are truncated to uint32.
 */
void or_rect_hist(const uint64_t *rows, uint64_t nrows, const uint64_t *cols, uint64_t ncols,
                  uint64_t *hist, unsigned nbins)
{
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(static)
        for (uint64_t r = 0; r < nrows; ++r) {
            const uint64_t w = rows[r];
            for (uint64_t s = 0; s < ncols; ++s) ++local[popc64(w ^ cols[s])];
        }
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            hist[b] += local[b];
        }
    }
}

/* Same as or_rect_hist with rows = Ap*i, i in [row_begin, row_end). */
void or_rect_hist_arith(uint64_t Ap, uint64_t row_begin, uint64_t row_end, int t32, const uint64_t *cols,
                        uint64_t ncols, uint64_t *hist, unsigned nbins)
{
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(static)
        for (uint64_t i = row_begin; i < row_end; ++i) {
            uint64_t w = Ap * i;
            if (t32) w &= 0xFFFFFFFFull;
            for (uint64_t s = 0; s < ncols; ++s) ++local[popc64(w ^ cols[s])];
        }
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            hist[b] += local[b];
        }
    }
}

/*
 * Host scaling of run_ancoding_grid (an_coding_grid.cu:316-322):
 * out[i] = (u128)(factor * hist[i]) in long double, factor = 2^n/m (1D) or
 * 4^n/(m*m2) (2D).  Bin 0 is overwritten by the scaled self-hit count.
 */
void or_grid_report(int gdim, uint64_t n, uint64_t m, uint64_t m2, const uint64_t *hist, unsigned count_counts,
                    u128 *out)
{
    long double factor = gdim == 1 ? powl(2.0L, (long double)n) / m
                                   : powl(4.0L, (long double)n) / (m * m2);
    out[0] = (u128)1 << n;
    for (unsigned i = 0; i < count_counts; ++i) out[i] = (u128)(factor * hist[i]);
}

/* ------------------------------------------------------------------------ */
/* Extended Hamming                                                           */
/* ------------------------------------------------------------------------ */

/*
 * Parity masks: the 32-bit masks of computeHamming32 (distance_distribution/src/hamming.h:54-64,
 * hamming.cu:43-57) are the low halves of the 64-bit patterns of
 * legacy/Hamming/ExtHamming64.hpp:31-37, so one table serves both; it is
 * truncated to the data width k.
 */
static const uint64_t OR_HMASK[7] = {
    0xAB55555556AAAD5Bull, 0xCD9999999B33366Dull, 0x78F1E1E1E3C3C78Eull, 0x01FE01FE03FC07F0ull,
    0x01FFFE0003FFF800ull, 0x01FFFFFFFC000000ull, 0xFE00000000000000ull};

/* Number of (non-overall) parity bits for k data bits: the masks that are
 * nonzero on the data (k=8:4, k=16/24:5, k=32:6, k=64:7; cf. hamming.cu:133,
 * where h = parity bits + 1). */
unsigned or_hamming_parity_bits(unsigned k)
{
    const uint64_t dm = k >= 64 ? ~0ull : ((1ull << k) - 1);
    unsigned r = 0;
    for (int t = 0; t < 7; ++t)
        if (OR_HMASK[t] & dm) ++r;
    return r;
}

/* Weight of the non-extended codeword: popcount(d) + sum of parity bits
 * (computeHamming<N>, hamming.cu:39-59; apply_w<N>, hamming.cpp:25-41). */
unsigned or_hamming_weight_nonext(uint64_t d, unsigned k)
{
    const uint64_t dm = k >= 64 ? ~0ull : ((1ull << k) - 1);
    d &= dm;
    unsigned w = popc64(d);
    for (int t = 0; t < 7; ++t)
        if (OR_HMASK[t] & dm) w += popc64(d & OR_HMASK[t]) & 1u;
    return w;
}

/* Full extended codeword as laid out by computeHamming08..32 (hamming.h:22-64):
 * data << h, parity bit t at position t+1, overall parity at bit 0.  Used
 * for brute-force sphere checks on small k. */
uint64_t or_hamming_codeword_ext(uint64_t d, unsigned k)
{
    const uint64_t dm = k >= 64 ? ~0ull : ((1ull << k) - 1);
    d &= dm;
    const unsigned r = or_hamming_parity_bits(k);
    uint64_t par = 0;
    unsigned t2 = 0;
    for (int t = 0; t < 7; ++t)
        if (OR_HMASK[t] & dm) {
            par |= (uint64_t)(popc64(d & OR_HMASK[t]) & 1u) << (t2 + 1);
            ++t2;
        }
    par |= (popc64(d) + popc64(par)) & 1u;
    return (d << (r + 1)) | par;
}

/* Non-extended weight histogram over data words d in [d_begin, d_end)
 * (hamming_cpu<N>, hamming.cpp:43-63 -- note the reference CPU loop starts
 * at d=1 while the GPU kernel dhamming (hamming.cu:64-81) starts at 0). */
void or_hamming_weights(unsigned k, uint64_t d_begin, uint64_t d_end, uint64_t *hist, unsigned nbins)
{
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(static)
        for (uint64_t d = d_begin; d < d_end; ++d) ++local[or_hamming_weight_nonext(d, k)];
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            hist[b] += local[b];
        }
    }
}

/* Weight histogram over an explicit list of data words. */
void or_hamming_weights_list(unsigned k, const uint64_t *d, uint64_t cnt, uint64_t *hist, unsigned nbins)
{
    memset(hist, 0, nbins * sizeof(uint64_t));
    for (uint64_t i = 0; i < cnt; ++i) ++hist[or_hamming_weight_nonext(d[i], k)];
}

/*
 * Post-processing of run_hamming (hamming.cu:182-208) from the non-extended
 * weight histogram h[]: out[0]=2^k, out[1]=N*2^k if spheres else 0, out[2]=0,
 * out[even i>=4] = h[i]+h[i-1] (odd weights gain the overall parity bit),
 * out[odd i>=3] = (i+1)*out[i+1] + (N-i+1)*out[i-1] if spheres else 0, and
 * out[i>=3] <<= k.  N = k + parity bits + 1.
 */
void or_hamming_report(unsigned k, const uint64_t *h, int with_1bit, u128 *out)
{
    const unsigned N = k + or_hamming_parity_bits(k) + 1;
    const unsigned cc = N + 1;
    for (unsigned i = 0; i < cc; ++i) out[i] = 0;
    out[0] = (u128)1 << k;
    out[1] = with_1bit ? ((u128)1 << k) * N : 0;
    out[2] = 0;
    for (unsigned i = 4; i < cc; i += 2) {
        out[i] = (u128)h[i] + (u128)h[i - 1];
        out[i - 1] = 0;
    }
    if (with_1bit)
        for (unsigned i = 3; i < cc; i += 2) {
            if (i + 1 < cc)
                out[i] = (u128)(i + 1) * out[i + 1] + (u128)(N - i + 1) * out[i - 1];
            else
                out[i] = (u128)(N - i + 1) * out[i - 1];
        }
    for (unsigned i = 3; i < cc; ++i) out[i] <<= k;
}

/*
 * Brute force over all ordered codeword pairs and the explicit 1-bit spheres
 * (hamming_native_short.cpp:52-77): counts[d(x,y)]++ and, with spheres,
 * counts[d(x, y ^ (1<<p))]++ for every bit p of the N-bit codeword.
 * O(4^k * N): small k only.
 */
void or_hamming_bruteforce(unsigned k, int with_1bit, u128 *out)
{
    const unsigned N = k + or_hamming_parity_bits(k) + 1;
    const uint64_t M = 1ull << k;
    uint64_t *cw = (uint64_t *)malloc(M * sizeof(uint64_t));
    for (uint64_t a = 0; a < M; ++a) cw[a] = or_hamming_codeword_ext(a, k);
    uint64_t tot[OR_MAXBINS];
    memset(tot, 0, sizeof(tot));
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(dynamic, 1)
        for (uint64_t a = 0; a < M; ++a) {
            const uint64_t x = cw[a];
            for (uint64_t b = 0; b < M; ++b) {
                const uint64_t y = cw[b];
                ++local[popc64(x ^ y)];
                if (with_1bit)
                    for (unsigned p = 0; p < N; ++p) ++local[popc64(x ^ (y ^ (1ull << p)))];
            }
        }
        for (unsigned b = 0; b <= N; ++b) {
#pragma omp atomic
            tot[b] += local[b];
        }
    }
    for (unsigned b = 0; b <= N; ++b) out[b] = tot[b];
    free(cw);
}

/*
 * Sample points of dhamming_grid_1d (hamming_grid.cu:53-74) with the host
 * set-up of run_hamming_grid (:122-156, single device): shard size S (the
 * Hamming::traits::Shards value for k), count_shards = m/S, sample
 * x = w + k'*step evaluated as (u64)((double)w + (double)k'*step) with
 * w = shard*S*(u64)step.  Returns the number of samples written.
 */
uint64_t or_hamming_grid_ref_samples(uint64_t m, double step, uint64_t S, uint64_t *out, uint64_t cap)
{
    const uint64_t nshards = m / S;
    uint64_t cnt = 0;
    for (uint64_t sh = 0; sh < nshards; ++sh) {
        const uint64_t w = sh * S * (uint64_t)step;
        for (uint64_t kk = 0; kk < S; ++kk) {
            uint64_t x = w;
            x = (uint64_t)((double)x + (double)kk * step);
            if (cnt < cap) out[cnt] = x;
            ++cnt;
        }
    }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
 * as easy as 1, 2, 3"), written from the published description: 10 rounds of
 * the 4x32 Philox S-box with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85.                                        */
/* ------------------------------------------------------------------------ */

void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t x0 = ctr_in[0], x1 = ctr_in[1], x2 = ctr_in[2], x3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        x0 = hi1 ^ x1 ^ k0;
        x1 = lo1;
        x2 = hi0 ^ x3 ^ k1;
        x3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

/*
 * Uniform data words for Monte Carlo, defined by this library (the reference
 * uses closed cuRAND Sobol32 streams, rand_gen.cu:140-168, which are replaced):
 * sample s of stream (seed, stream) is taken from the Philox block
 * ctr = (s / per, (s / per) >> 32, stream, 0x43520001), key = (seed, seed >> 32);
 * for k <= 32 each block yields per=4 words out[s%4] & (2^k-1); for k > 32
 * per=2 words (out[2j] | out[2j+1] << 32) & (2^k-1).
 */
uint64_t or_mc_word(uint64_t seed, uint32_t stream, uint64_t s, unsigned k)
{
    const unsigned per = k <= 32 ? 4 : 2;
    const uint64_t blk = s / per;
    const uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0x43520001u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    const uint64_t dm = k >= 64 ? ~0ull : ((1ull << k) - 1);
    if (per == 4) return (uint64_t)o[s % 4] & dm;
    const unsigned j = (unsigned)(s % 2);
    return ((uint64_t)o[2 * j] | ((uint64_t)o[2 * j + 1] << 32)) & dm;
}

/* Monte Carlo non-extended weight histogram over samples [s_begin, s_end)
 * (replaces dhamming_mc, hamming_mc.cu:57-81). */
void or_hamming_mc(unsigned k, uint64_t seed, uint32_t stream, uint64_t s_begin, uint64_t s_end, uint64_t *hist,
                   unsigned nbins)
{
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(static)
        for (uint64_t s = s_begin; s < s_end; ++s)
            ++local[or_hamming_weight_nonext(or_mc_word(seed, stream, s, k), k)];
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            hist[b] += local[b];
        }
    }
}

/*
 * Sample j of row i of the AN Monte Carlo (this library's definition; the
 * reference draws (T)(2^n * U) from closed cuRAND streams, an_coding_mc.cu:125):
 * Philox block b of counter (b, b >> 32, i mod 2^32, 0x43520001 ^ (i >> 32)),
 * key (seed, seed >> 32); n <= 32: b = j/4, word j%4; n > 32: b = j/2, the
 * 64-bit word out[2(j%2)] | out[2(j%2)+1] << 32; masked to n bits.
 */
uint64_t or_an_mc_sample(uint64_t seed, uint64_t i, uint64_t j, unsigned n)
{
    const unsigned per = n <= 32 ? 4 : 2;
    const uint64_t blk = j / per;
    const uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)i, 0x43520001u ^ (uint32_t)(i >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    const uint64_t dm = n >= 64 ? ~0ull : ((1ull << n) - 1);
    if (per == 4) return (uint64_t)o[j % 4] & dm;
    const unsigned h = (unsigned)(j % 2);
    return ((uint64_t)o[2 * h] | ((uint64_t)o[2 * h + 1] << 32)) & dm;
}

/*
 * Monte Carlo AN histogram (replaces dancoding_mc_shared,
 * an_coding_mc.cu:94-158): every word i in [i_begin, i_end) of [0, 2^n) is
 * compared with the m codewords A*u_ij, u_ij = or_an_mc_sample(seed, i, j, n).
 * hist is the raw (word, sample) ordered-pair histogram.
 */
void or_an_mc_rows(uint64_t A, unsigned n, uint64_t m, uint64_t seed, uint64_t i_begin, uint64_t i_end,
                   uint64_t *hist, unsigned nbins)
{
#pragma omp parallel
    {
        uint64_t local[OR_MAXBINS];
        memset(local, 0, sizeof(local));
#pragma omp for schedule(static)
        for (uint64_t i = i_begin; i < i_end; ++i) {
            const uint64_t w = A * i;
            for (uint64_t j = 0; j < m; ++j) ++local[popc64(w ^ (A * or_an_mc_sample(seed, i, j, n)))];
        }
        for (unsigned b = 0; b < nbins; ++b) {
#pragma omp atomic
            hist[b] += local[b];
        }
    }
}

void or_an_mc(uint64_t A, unsigned n, uint64_t m, uint64_t seed, uint64_t *hist, unsigned nbins)
{
    or_an_mc_rows(A, n, m, seed, 0, 1ull << n, hist, nbins);
}

/* ------------------------------------------------------------------------ */
/* ANcandidates                                                               */
/* ------------------------------------------------------------------------ */

/*
 * min over d in [1, 2^k) of popcount(c ^ 3c), c = A*d, the length of the
 * shortest signed-digit (NAF) representation (worker, ANcandidates/ANcandidates.cpp:35-160,
 * without the early break, which only prunes A's that cannot win).
 */
uint64_t or_an_min_naf(uint64_t A, unsigned k)
{
    uint64_t best = ~0ull;
    const uint64_t maxc = ((1ull << k) - 1) * A;
    for (uint64_t c = A; c <= maxc; c += A) {
        const uint64_t p = popc64(c ^ (3ull * c));
        if (p < best) best = p;
    }
    return best;
}

/* For all odd A with exactly abits bits: best = max_A min_naf(A), and the
 * A's attaining it in increasing order (ANcandidates.cpp:254-299).
 * Returns the number of candidates (at most cap written). */
uint64_t or_an_candidates(unsigned k, unsigned abits, uint64_t *best_out, uint64_t *cands, uint64_t cap)
{
    if (abits < 2 || k < 1) {
        *best_out = 0;
        return 0;
    }
    const uint64_t minA = (1ull << (abits - 1)) + 1, maxA = (1ull << abits) - 1;
    const uint64_t numA = (maxA - minA) / 2 + 1;
    uint64_t *vals = (uint64_t *)malloc(numA * sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t t = 0; t < numA; ++t) vals[t] = or_an_min_naf(minA + 2 * t, k);
    uint64_t best = 0;
    for (uint64_t t = 0; t < numA; ++t)
        if (vals[t] > best) best = vals[t];
    uint64_t cnt = 0;
    for (uint64_t t = 0; t < numA; ++t)
        if (vals[t] == best) {
            if (cnt < cap) cands[cnt] = minA + 2 * t;
            ++cnt;
        }
    free(vals);
    *best_out = best;
    return cnt;
}

int or_num_threads(void) { return omp_get_max_threads(); }
