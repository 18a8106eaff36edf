// Extended-Hamming weight histograms and AN Monte Carlo pair histograms (gfx950).
//
// Hamming: by linearity the distance distribution of the code is the weight
// distribution of its codewords (the reference's method,
// distance_distribution/src/hamming.cu:183-208), so the kernel histograms the
// weight of the non-extended codeword, popcount(d) + sum_t parity(d & mask_t)
// (computeHamming<N>, hamming.cu:39-59), over a stream of data words d:
//   range   d = first + s                       (exhaustive: dhamming, hamming.cu:64-81)
//   stride  d = (first + s*stride) mod 2^k      (integer grid)
//   list    d = list[s]                         (the reference grid's sample points)
//   mc      d = Philox word s of (seed, stream) (replaces dhamming_mc's cuRAND draw)
// AN Monte Carlo: every word A*i against m Philox samples A*u_ij
// (replaces dancoding_mc_shared, an_coding_mc.cu:94-158).
#include <algorithm>
#include <type_traits>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

using namespace dev;

// Work item = 64 lanes x job.lane_samples samples.  Items of up to 2^18 words
// keep the work-queue counter (one device-scope atomic per item) below its
// throughput limit (about 8e7 items/s, reached with 4096-word items at 3.4e11
// words/s and with 32768-word items at 2.5e12 words/s on the exhaustive fast
// path); small jobs use smaller items so that they still fill the device.

// Syndrome tables: the syndrome (bit t = parity(d & m_t)) is linear over
// GF(2), so it is the XOR of the syndromes of the data word's bit fields.
// Each 32-bit half of the word is cut into fields of 7, 7, 7, 7 and 4 bits;
// table f (f = 5h + i for field i of half h) holds the syndromes of the
// field's 2^7 (or 2^4) values, one byte each, at byte offset 128 f: 5 tables
// for 32-bit data, 10 for 64-bit.  A 128-byte table spans 32 consecutive
// dwords, one in each LDS bank, so two lanes of a lookup only meet in a bank
// on the same dword, whose read they share: the lookups are free of bank
// conflicts (a 256-entry byte table puts two dwords in every bank).  The masks
// are truncated to k bits, so bits of d above k do not change the syndrome.
constexpr uint32_t kFields = 5;  // fields per 32-bit half

template <bool W64>
constexpr uint32_t kTabBytes = (W64 ? 2 : 1) * kFields * 128;

template <bool W64>
__device__ void build_tables(uint8_t *tab, const uint64_t *masks) {
    uint64_t m[7];
#pragma unroll
    for (int t = 0; t < 7; ++t) m[t] = masks[t];
    for (uint32_t e = threadIdx.x; e < kTabBytes<W64>; e += 64) {
        const uint32_t f = e >> 7, v = e & 127, h = f / kFields, i = f % kFields;
        uint32_t syn = 0;
        if (i < kFields - 1 || v < 16) {
            const uint64_t x = (uint64_t)v << (32 * h + 7 * i);
#pragma unroll
            for (int t = 0; t < 7; ++t) syn |= (uint32_t)(__popcll(x & m[t]) & 1) << t;
        }
        tab[e] = (uint8_t)syn;
    }
    __syncthreads();
}

// The table lookups of U words (all issued before any is combined, so they
// are in flight together) and their combination into weights: popcount of the
// data word plus popcount of the syndrome, the weight of the non-extended
// codeword.
template <bool W64, int U>
__device__ __forceinline__ void table_reads(const uint64_t (&d)[U], const uint8_t *tab,
                                            uint32_t (&r)[U][(W64 ? 2 : 1) * kFields]) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (uint32_t h = 0; h < (W64 ? 2u : 1u); ++h) {
            const uint32_t x = (uint32_t)(d[u] >> (32 * h));
            const uint8_t *t = tab + 128 * kFields * h;
            r[u][kFields * h + 0] = t[x & 127];
            r[u][kFields * h + 1] = t[128 + ((x >> 7) & 127)];
            r[u][kFields * h + 2] = t[256 + ((x >> 14) & 127)];
            r[u][kFields * h + 3] = t[384 + ((x >> 21) & 127)];
            r[u][kFields * h + 4] = t[512 + (x >> 28)];
        }
}

// FULL: dmask covers the whole word (k = 32 or 64), so the AND is left out.
template <bool W64, int U, bool FULL = false>
__device__ __forceinline__ void table_weights(const uint64_t (&d)[U], uint64_t dmask,
                                              const uint32_t (&r)[U][(W64 ? 2 : 1) * kFields], uint32_t (&w)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        uint32_t syn = r[u][0];
#pragma unroll
        for (uint32_t f = 1; f < (W64 ? 2 : 1) * kFields; ++f) syn ^= r[u][f];
        const uint64_t dm = FULL ? d[u] : d[u] & dmask;
        w[u] = __popc(syn) + (W64 ? (uint32_t)__popcll(dm) : (uint32_t)__popc((uint32_t)dm));
    }
}

template <bool W64, int U>
__device__ __forceinline__ void weights_tab(const uint64_t (&d)[U], uint64_t dmask, const uint8_t *tab,
                                            uint32_t (&w)[U]) {
    uint32_t r[U][(W64 ? 2 : 1) * kFields];
    table_reads<W64, U>(d, tab, r);
    table_weights<W64, U>(d, dmask, r, w);
}

// Syndrome of a data word: bit t = parity(x & m_t) (the seventh mask is zero
// for k <= 32).
__device__ __forceinline__ uint32_t syndrome32(uint32_t x, const uint64_t (&m)[7]) {
    uint32_t syn = 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) syn |= (__popc(x & (uint32_t)m[t]) & 1u) << t;
    return syn;
}
__device__ __forceinline__ uint32_t syndrome64(uint64_t x, const uint64_t (&m)[7]) {
    uint32_t syn = 0;
#pragma unroll
    for (int t = 0; t < 7; ++t) syn |= (uint32_t)(__popcll(x & m[t]) & 1) << t;
    return syn;
}

// Exhaustive fast path: an aligned item holds the words d = dh | 64*q | lane
// (disjoint bit fields), and both popcount and syndrome split over disjoint
// fields (the syndrome is linear over GF(2)): weight(d) = popc(dh) + popc(x) +
// popc(S(dh) ^ S(x)), x = 64*ql | lane, with q = 16*qh + ql.  S(x) and
// popc(x) are per-lane constants for ql < 16; S(dh), popc(dh) are wave-uniform.
constexpr uint32_t kFastUnroll = 16;

// One Monte Carlo item: lane l takes the Philox blocks B0 + 64q + l, q <
// spl/PER, B0 = (first + base)/PER (PER = 4 words per block for k <= 32, 2
// for k > 32: sample s of the job is word s mod PER of block s/PER).  Three
// blocks are in flight: at step q the table lookups of block q+1 are issued,
// then the Philox rounds of block q+2 run interleaved with the combination
// of block q, whose lookups were issued a step earlier.  The last two steps
// generate and look up blocks past the item's end, which are not counted.
// CHECK: the item reaches past the job's sample count, so samples >= count
// go to the dummy row.
template <uint32_t PER>
__device__ __forceinline__ void mc_words(const uint32_t (&c)[4], uint64_t (&x)[PER]) {
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u)
        x[u] = PER == 4 ? (uint64_t)c[u] : ((uint64_t)c[2 * u] | ((uint64_t)c[2 * u + 1] << 32));
}

__device__ __forceinline__ void mc_block(uint32_t (&c)[4], uint64_t b, uint32_t stream, uint32_t k0, uint32_t k1) {
    c[0] = (uint32_t)b;
    c[1] = (uint32_t)(b >> 32);
    c[2] = stream;
    c[3] = kMcTag;
    philox4x32_10(c, k0, k1);
}

template <bool W64, uint32_t PER, bool CHECK, bool FULL>
__device__ __forceinline__ void mc_blocks(uint32_t *lane_col, const uint8_t *tab, const HamJob &job, uint64_t base,
                                          uint32_t spl, uint64_t count, uint64_t dmask, uint32_t dummy) {
    constexpr uint32_t F = (W64 ? 2 : 1) * kFields;
    const uint32_t lane = threadIdx.x;
    const uint32_t k0 = (uint32_t)job.seed, k1 = (uint32_t)(job.seed >> 32);
    const uint32_t nq = spl / PER;
    const uint64_t b0 = (job.first + base) / PER + lane;
    uint32_t c[4];
    uint64_t x[2][PER];
    uint32_t r[2][PER][F];
    mc_block(c, b0, job.stream, k0, k1);
    mc_words<PER>(c, x[0]);
    table_reads<W64, PER>(x[0], tab, r[0]);
    mc_block(c, b0 + 64, job.stream, k0, k1);
    // nq is even (spl is a power of two >= 16): two steps per iteration, the
    // buffers of blocks q and q+1 trading places
    for (uint32_t q = 0; q < nq; q += 2) {
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            mc_words<PER>(c, x[h ^ 1]);
            table_reads<W64, PER>(x[h ^ 1], tab, r[h ^ 1]);
            mc_block(c, b0 + 64ull * (q + h + 2), job.stream, k0, k1);
            uint32_t w[PER];
            table_weights<W64, PER, FULL>(x[h], dmask, r[h], w);
            const uint64_t s = base + ((uint64_t)(q + h) * 64 + lane) * PER;  // first local sample of this lane
#pragma unroll
            for (uint32_t u = 0; u < PER; ++u)
                lds_inc(lane_col + ((!CHECK || s + u < count) ? w[u] : dummy) * kStride);
        }
    }
}

template <bool W64, bool CHECK>
__device__ __forceinline__ void mc_item(uint32_t *lane_col, const uint8_t *tab, const HamJob &job, uint64_t base,
                                        uint32_t spl, uint64_t count, uint64_t dmask, uint32_t dummy) {
    if (!W64 || job.k <= 32) {
        if (job.k == 32)
            mc_blocks<W64, 4, CHECK, true>(lane_col, tab, job, base, spl, count, dmask, dummy);
        else
            mc_blocks<W64, 4, CHECK, false>(lane_col, tab, job, base, spl, count, dmask, dummy);
    } else if (job.k == 64) {
        mc_blocks<W64, 2, CHECK, true>(lane_col, tab, job, base, spl, count, dmask, dummy);
    } else {
        mc_blocks<W64, 2, CHECK, false>(lane_col, tab, job, base, spl, count, dmask, dummy);
    }
}

// Range, stride and list items off the exhaustive fast path.  Lane l takes
// the words d_q = first + (base + 64q + l) * stride for q < spl (stride 1 for
// a range), so d advances by 64*stride per step (modulo 2^64, or 2^32 for
// k <= 32: only the low k bits reach the weight), or list[base + 64q + l].
// kSeqUnroll steps per iteration, all table reads (and list loads) of the
// batch issued before its increments, so the reads overlap instead of each
// waiting behind the previous increment (8 steps of 5 lookups for 32-bit
// words, 4 of 10 for 64-bit).  spl is a multiple of kSeqUnroll.
// CHECK: steps past the job's count go to the dummy row.
template <bool W64>
constexpr uint32_t kSeqUnroll = W64 ? 4 : 8;

template <bool W64, bool CHECK, bool LIST>
__device__ __forceinline__ void seq_item(uint32_t *lane_col, const uint8_t *tab, const HamJob &job, uint64_t base,
                                         uint32_t spl, uint64_t count, uint64_t dmask, uint32_t dummy) {
    using word_t = typename std::conditional<W64, uint64_t, uint32_t>::type;
    const uint32_t lane = threadIdx.x;
    const uint64_t rem = count - base;
    const uint32_t left = rem < 64ull * spl ? (uint32_t)rem : 64u * spl;  // valid samples in the item
    const uint64_t stride = job.source == kHamRange ? 1 : job.stride;
    word_t d = (word_t)(job.first + (base + lane) * stride);
    const word_t step = (word_t)(64 * stride);
    const uint64_t *list = LIST ? job.list + base + lane : nullptr;
    constexpr uint32_t U = kSeqUnroll<W64>;
    for (uint32_t q = 0; q < spl; q += U) {
        uint64_t x[U];
        uint32_t w[U];
        if (LIST) {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                x[u] = (!CHECK || (q + u) * 64 + lane < left) ? list[(q + u) * 64] : 0;
        } else {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                x[u] = (uint64_t)d;
                d += step;
            }
        }
        weights_tab<W64, U>(x, dmask, tab, w);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            lds_inc(lane_col + ((!CHECK || (q + u) * 64 + lane < left) ? w[u] : dummy) * kStride);
    }
}

__device__ __forceinline__ unsigned long long *copy_out(const HamJob &job) {
    return hist_copy(job.out, job.nbins, job.out_copies);
}

__device__ __forceinline__ uint32_t find_job(const HamJob *jobs, uint32_t njobs, uint64_t t, uint32_t cur) {
    if (cur != 0xFFFFFFFFu && t >= jobs[cur].item_begin && t < jobs[cur].item_begin + jobs[cur].items) return cur;
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].item_begin <= t) lo = mid; else hi = mid - 1;
    }
    return uniform32(lo);
}

template <bool W64>
__global__ __launch_bounds__(64) void ham_kernel(const HamJob *__restrict__ jobs, uint32_t njobs, uint64_t total_items,
                                                 unsigned long long *__restrict__ counter, uint32_t lds_rows) {
    // the tables at a fixed LDS address (their offsets fold into the reads),
    // the histogram rows in the dynamic LDS after them
    __shared__ uint8_t tab[kTabBytes<W64>];
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, lds_rows);
    uint32_t *const lane_col = lds + lane;
    const uint32_t dummy = lds_rows - 1;
    uint32_t cur = 0xFFFFFFFFu;
    uint64_t pending = 0;

    for (uint64_t t = first_item(); t < total_items; t = following_item(counter, t)) {
        const uint32_t jb = find_job(jobs, njobs, t, cur);
        const HamJob &job = jobs[jb];
        const uint32_t spl = job.lane_samples;
        if (jb != cur || pending + spl >= 0xFFFFFFFFull) {
            if (cur != 0xFFFFFFFFu) flush_rows(lds, 0, jobs[cur].nbins, copy_out(jobs[cur]));
            if (jb != cur) build_tables<W64>(tab, job.masks);
            cur = jb;
            pending = 0;
        }
        pending += spl;
        uint64_t m[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) m[i] = job.masks[i];
        const uint64_t dmask = job.k >= 64 ? ~0ull : ((1ull << job.k) - 1);
        const uint64_t base = (t - job.item_begin + job.item_offset) * (64ull * spl);  // first sample of the item
        const uint64_t count = job.count;

        if (job.source == kHamMc) {
            if (base + 64ull * spl <= count)
                mc_item<W64, false>(lane_col, tab, job, base, spl, count, dmask, dummy);
            else
                mc_item<W64, true>(lane_col, tab, job, base, spl, count, dmask, dummy);
        } else if (job.source == kHamRange && base + 64ull * spl <= count && spl >= kFastUnroll &&
                   (spl & (spl - 1)) == 0 && ((job.first + base) & (64ull * spl - 1)) == 0 &&
                   64ull * spl - 1 <= dmask) {
            // full item of consecutive words (exhaustive) whose first word is a
            // multiple of the item size 64*spl, a power of two: the item's
            // words are d0 | x for x < 64*spl (disjoint bits)
            // per-lane constants of the item, S(x) from the tables (x < 2^10
            // spans fields 0 and 1); kept to this path so that they hold no
            // registers in the others
            uint32_t fsyn[kFastUnroll], fbase[kFastUnroll];
#pragma unroll
            for (uint32_t ql = 0; ql < kFastUnroll; ++ql) {
                const uint32_t x = (ql << 6) | lane;
                fsyn[ql] = tab[x & 127] ^ tab[128 + (x >> 7)];
                fbase[ql] = lane * 4 + ((uint32_t)__popc(x) << 8);  // byte offset in the LDS rows
            }
            const uint64_t d0 = (job.first + base) & dmask;
            for (uint32_t qh = 0; qh < spl / kFastUnroll; ++qh) {
                uint32_t sh, ph;
                if (W64) {
                    const uint64_t dh = uniform64(d0 | ((uint64_t)qh << 10));
                    sh = uniform32(syndrome64(dh, m));
                    ph = uniform32((uint32_t)__popcll(dh));
                } else {
                    const uint32_t dh = uniform32((uint32_t)d0 | (qh << 10));
                    sh = uniform32(syndrome32(dh, m));
                    ph = uniform32(__popc(dh));
                }
#pragma unroll
                for (uint32_t ql = 0; ql < kFastUnroll; ++ql) {
                    const uint32_t w = __popc(fsyn[ql] ^ sh) + ph;
                    lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(lds) + ((w << 8) + fbase[ql])));
                }
            }
        } else {
            const bool full = base + 64ull * spl <= count;
            if (job.source == kHamList) {
                if (full)
                    seq_item<W64, false, true>(lane_col, tab, job, base, spl, count, dmask, dummy);
                else
                    seq_item<W64, true, true>(lane_col, tab, job, base, spl, count, dmask, dummy);
            } else if (full) {
                seq_item<W64, false, false>(lane_col, tab, job, base, spl, count, dmask, dummy);
            } else {
                seq_item<W64, true, false>(lane_col, tab, job, base, spl, count, dmask, dummy);
            }
        }
    }
    if (cur != 0xFFFFFFFFu) flush_rows(lds, 0, jobs[cur].nbins, copy_out(jobs[cur]));
}

// AN Monte Carlo: item = (row block of 64*RM words, sample chunk [j0,
// j0+CJ)).  Row i draws u_ij from the Philox block b of counter
// (b, b >> 32, i mod 2^32, tag ^ (i >> 32)), key (seed, seed >> 32):
//   n <= 32: b = j/4, word j%4, masked to n bits;
//   n >  32: b = j/2, words 2(j%2) (low) and 2(j%2)+1 (high), masked to n bits;
// the codewords A*i and A*u_ij are compared.
constexpr int RM = 4;
constexpr uint32_t CJ = 256;  // samples per item (multiple of 4)

// One AN Monte Carlo item: samples [jlo, j1) of the chunk that starts at j0
// (a multiple of CJ, so of the words per Philox block).  W1: every codeword
// fits 32 bits (32-bit multiply, xor and popcount).  S64: 64-bit samples
// (n > 32).  CHECK: rows past 2^n and samples outside [jlo, j1) (a partial
// chunk) go to the dummy row; without CHECK, jlo = j0 and j1 - j0 is a
// multiple of the words per Philox block.
template <bool W1, bool S64, bool CHECK>
__device__ __forceinline__ void an_mc_item(uint32_t *lane_col, uint64_t A, uint64_t rb, uint64_t nwords, uint64_t j0,
                                           uint64_t jlo, uint64_t j1, uint64_t smask, uint32_t k0, uint32_t k1,
                                           uint32_t dummy) {
    const uint32_t lane = threadIdx.x;
    constexpr uint32_t per = S64 ? 2 : 4;
#pragma unroll
    for (int r = 0; r < RM; ++r) {
        const uint64_t i = rb * (64ull * RM) + (uint64_t)r * 64 + lane;
        const bool ok = !CHECK || i < nwords;
        const uint64_t w = A * i;
        const uint32_t tag = S64 ? kMcTag ^ (uint32_t)(i >> 32) : kMcTag;
        for (uint64_t j = CHECK ? jlo / per * per : j0; j < j1; j += per) {
            const uint64_t blk = j / per;
            uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)i, tag};
            philox4x32_10(c, k0, k1);
            if (S64) {
#pragma unroll
                for (uint32_t u = 0; u < 2; ++u) {
                    const uint64_t x = ((uint64_t)c[2 * u] | ((uint64_t)c[2 * u + 1] << 32)) & smask;
                    const uint32_t d = (uint32_t)__popcll(w ^ (A * x));
                    lds_inc(lane_col + ((!CHECK || (ok && j + u >= jlo && j + u < j1)) ? d : dummy) * kStride);
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const uint32_t x = c[u] & (uint32_t)smask;
                    const uint32_t d = W1 ? (uint32_t)__popc((uint32_t)w ^ ((uint32_t)A * x))
                                          : (uint32_t)__popcll(w ^ (A * (uint64_t)x));
                    lds_inc(lane_col + ((!CHECK || (ok && j + u >= jlo && j + u < j1)) ? d : dummy) * kStride);
                }
            }
        }
    }
}

// Item t of segment s (McSegment: samples [j0, j1) with a histogram of its
// own) = (row group of G row blocks, sample chunk): local index t -
// item_begin = rg * chunks + c, chunk c covering [a + c*CJ, a + (c+1)*CJ) of
// [j0, j1), a = j0 rounded down to CJ.  G is chosen on the host so that an
// item holds about 2^18 pairs (one work-queue step per item).  Segments are
// laid out one after the other, so a wave's items stay in one segment for
// long runs; its LDS rows go to the segment's histogram when the segment
// changes.
__global__ __launch_bounds__(64) void an_mc_kernel(uint64_t A, uint32_t n, uint64_t seed, uint32_t G,
                                                   const McSegment *__restrict__ segs, uint32_t nsegs,
                                                   uint64_t total_items, unsigned long long *__restrict__ counter,
                                                   uint32_t nbins, uint32_t lds_rows) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x;
    lds_clear(lds, lds_rows);
    uint32_t *const lane_col = lds + lane;
    const uint32_t dummy = lds_rows - 1;
    const uint64_t nwords = 1ull << n;
    const uint64_t RBlocks = (nwords + 64ull * RM - 1) / (64ull * RM);
    const uint64_t smask = nwords - 1;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    // all codewords (A * (2^n - 1)) below 2^32
    const bool w1 = ((unsigned __int128)A * (nwords - 1)) >> 32 == 0;
    const bool s64 = n > 32;
    uint64_t pending = 0;  // increments per lane counter since the last flush (bound)
    uint32_t cur = 0xFFFFFFFFu;  // segment whose counts the LDS rows hold
    for (uint64_t t = first_item(); t < total_items; t = following_item(counter, t)) {
        uint32_t sg = cur == 0xFFFFFFFFu ? 0 : cur;
        while (sg + 1 < nsegs && t >= segs[sg + 1].item_begin) ++sg;
        while (t < segs[sg].item_begin) --sg;
        sg = uniform32(sg);
        const McSegment &seg = segs[sg];
        const uint64_t local = t - seg.item_begin;
        const uint64_t rg = uniform64(local / seg.chunks), c = uniform64(local - rg * seg.chunks);
        const uint64_t j0 = (seg.j0 / CJ + c) * CJ;
        const uint64_t jlo = j0 > seg.j0 ? j0 : seg.j0;
        const uint64_t j1 = j0 + CJ < seg.j1 ? j0 + CJ : seg.j1;
        const uint64_t add = (uint64_t)G * RM * (j1 - j0);
        if (sg != cur || pending + add >= 0xFFFFFFFFull) {
            if (cur != 0xFFFFFFFFu) flush_rows(lds, 0, nbins, segs[cur].out);
            cur = sg;
            pending = 0;
        }
        pending += add;
        for (uint32_t g = 0; g < G; ++g) {
            const uint64_t rb = rg * G + g;
            if (rb >= RBlocks) break;
            // whole rows and whole Philox blocks: no masking
            const bool full = (rb + 1) * (64ull * RM) <= nwords && jlo == j0 && (j1 - j0) % (s64 ? 2 : 4) == 0;
            if (s64) {
                if (full)
                    an_mc_item<false, true, false>(lane_col, A, rb, nwords, j0, jlo, j1, smask, k0, k1, dummy);
                else
                    an_mc_item<false, true, true>(lane_col, A, rb, nwords, j0, jlo, j1, smask, k0, k1, dummy);
            } else if (w1 && full) {
                an_mc_item<true, false, false>(lane_col, A, rb, nwords, j0, jlo, j1, smask, k0, k1, dummy);
            } else if (full) {
                an_mc_item<false, false, false>(lane_col, A, rb, nwords, j0, jlo, j1, smask, k0, k1, dummy);
            } else {
                an_mc_item<false, false, true>(lane_col, A, rb, nwords, j0, jlo, j1, smask, k0, k1, dummy);
            }
        }
    }
    if (cur != 0xFFFFFFFFu) flush_rows(lds, 0, nbins, segs[cur].out);
}

// Dynamic LDS of a Hamming wave: the histogram rows (the syndrome tables are
// the kernel's static LDS).
size_t ham_lds_bytes(bool, uint32_t lds_rows) { return (size_t)lds_rows * kStride * sizeof(uint32_t); }

}  // namespace

cr_status launch_hamming(const HamJob *d_jobs, uint32_t njobs, uint64_t total_items, bool w64,
                         unsigned long long *counter, bool allow_static, uint32_t lds_rows, int cus,
                         hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
    const void *k = w64 ? reinterpret_cast<const void *>(&ham_kernel<true>)
                        : reinterpret_cast<const void *>(&ham_kernel<false>);
    if (lds_rows < 2 || lds_rows > 131) return fail(CR_ERR_INVALID_ARGUMENT, "bad histogram rows %u", lds_rows);
    const size_t lds = ham_lds_bytes(w64, lds_rows);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, lds) != hipSuccess || per_cu <= 0)
        return fail(CR_ERR_DEVICE, "occupancy query failed for the Hamming kernel");
    uint64_t waves = (uint64_t)per_cu * cus;
    if (waves > total_items) waves = total_items;
    if (waves == 0) waves = 1;
    if (allow_static && total_items > 0) {
        // static assignment over at most one wave per CU less than the
        // resident capacity, each wave taking the same number of items
        counter = nullptr;
        const uint64_t cap = (uint64_t)(per_cu > 1 ? per_cu - 1 : 1) * cus;
        if (waves > cap) waves = cap;
        const uint64_t per = (total_items + waves - 1) / waves;
        waves = (total_items + per - 1) / per;
    }
    if (counter) CR_TRY(fill_zero_u64(counter, 1, stream));
    if (ev0) CR_HIP(hipEventRecord(ev0, stream));
    const HamJob *a_jobs = d_jobs;
    uint32_t a_nj = njobs, a_rows = lds_rows;
    uint64_t a_items = total_items;
    unsigned long long *a_counter = counter;
    void *args[] = {(void *)&a_jobs, &a_nj, &a_items, &a_counter, &a_rows};
    CR_HIP(hipLaunchKernel(k, dim3((uint32_t)waves), dim3(64), args, lds, stream));
    if (ev1) CR_HIP(hipEventRecord(ev1, stream));
    return CR_OK;
}

uint64_t hamming_items(uint64_t count, uint32_t lane_samples) {
    const uint64_t per_item = 64ull * lane_samples;
    return (count + per_item - 1) / per_item;
}

uint32_t hamming_waves_per_cu(uint32_t lds_rows) {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void *>(&ham_kernel<false>), 64,
                                                     ham_lds_bytes(false, lds_rows)) != hipSuccess)
        a = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void *>(&ham_kernel<true>), 64,
                                                     ham_lds_bytes(true, lds_rows)) != hipSuccess)
        b = 0;
    return (uint32_t)std::max(std::max(a, b), 1);
}

cr_status launch_an_mc(uint64_t A, uint32_t n, uint64_t seed, McSegment *segs, uint32_t nsegs, McSegment *d_segs,
                       unsigned long long *counter, uint32_t nbins, int cus, hipStream_t stream, hipEvent_t ev0,
                       hipEvent_t ev1) {
    const uint32_t lds_rows = nbins + 1;
    if (lds_rows > 66) return fail(CR_ERR_INVALID_ARGUMENT, "bad histogram rows %u", lds_rows);
    if (nsegs == 0) return fail(CR_ERR_INVALID_ARGUMENT, "no sample range");
    const void *k = reinterpret_cast<const void *>(&an_mc_kernel);
    const size_t lds = (size_t)lds_rows * kStride * sizeof(uint32_t);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, lds) != hipSuccess || per_cu <= 0)
        return fail(CR_ERR_DEVICE, "occupancy query failed for the AN Monte Carlo kernel");
    const uint64_t RBlocks = ((1ull << n) + 64ull * RM - 1) / (64ull * RM);
    uint64_t total_chunks = 0, longest = 0;
    for (uint32_t s = 0; s < nsegs; ++s) {
        if (segs[s].j1 <= segs[s].j0) return fail(CR_ERR_INVALID_ARGUMENT, "empty sample range");
        segs[s].chunks = (segs[s].j1 - segs[s].j0 / CJ * CJ + CJ - 1) / CJ;
        total_chunks += segs[s].chunks;
        longest = std::max(longest, segs[s].j1 - segs[s].j0);
    }
    const uint64_t resident = (uint64_t)per_cu * cus;
    // row blocks per item: about 2^18 pairs per item (a power of two up to
    // 4096), keeping at least 16 items per resident wave
    const uint64_t chunk = longest < CJ ? longest : CJ;
    uint32_t G = 1;
    while (G < 4096 && (uint64_t)G * 64 * RM * chunk < (1ull << 18) &&
           ((RBlocks + 2 * G - 1) / (2 * G)) * total_chunks >= 16 * resident)
        G *= 2;
    const uint64_t RG = (RBlocks + G - 1) / G;
    uint64_t items = 0;
    for (uint32_t s = 0; s < nsegs; ++s) {
        segs[s].item_begin = items;
        items += RG * segs[s].chunks;
    }
    uint64_t waves = resident;
    if (waves > items) waves = items;
    CR_HIP(hipMemcpyAsync(d_segs, segs, nsegs * sizeof(McSegment), hipMemcpyHostToDevice, stream));
    CR_TRY(fill_zero_u64(counter, 1, stream));
    if (ev0) CR_HIP(hipEventRecord(ev0, stream));
    uint64_t a_A = A, a_seed = seed, a_items = items;
    uint32_t a_n = n, a_G = G, a_nsegs = nsegs, a_nbins = nbins, a_rows = lds_rows;
    const McSegment *a_segs = d_segs;
    unsigned long long *a_counter = counter;
    void *args[] = {&a_A, &a_n, &a_seed, &a_G, &a_segs, &a_nsegs, &a_items, &a_counter, &a_nbins, &a_rows};
    CR_HIP(hipLaunchKernel(k, dim3((uint32_t)waves), dim3(64), args, lds, stream));
    if (ev1) CR_HIP(hipEventRecord(ev1, stream));
    return CR_OK;
}

}  // namespace cr
