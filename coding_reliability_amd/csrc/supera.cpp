// Super-A searches over many A with one batched launch for all exhaustive
// histograms, followed by the reference drivers' sequential bookkeeping.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "coding_reliability.h"
#include "cr_internal.h"
#include "cr_reference_api.h"

namespace cr {
void supera_row(std::stringstream &ss, uintll_t A, uint_t h, const Flags &flags, double *times, uint128_t *counts,
                unsigned *id);
}

using namespace cr;

namespace {

uint_t ceil_h(uintll_t A) { return (uint_t)ceil(log((double)A) / log(2.0)); }
uint_t floor_h(uintll_t A) { return (uint_t)(floor(log((double)A) / log(2.0)) + 1); }

// first nonzero count below `limit` (an_coding.cu:295-308, super_a/src/main.cpp:157-170)
void lowest(const uint128_t *counts, uint_t limit, uintll_t &minb, uint128_t &mincb) {
    minb = 0xFFFF;
    mincb = static_cast<uint128_t>(-1);
    for (uint_t i = 1; i < limit; ++i)
        if (counts[i] != 0 && counts[i] < mincb) {
            minb = i;
            mincb = counts[i];
            break;
        }
}

struct CodeResult {
    uint128_t counts[COUNTS_MAX_WIDTH + 1] = {0};
    double kernel_s = 0, total_s = 0;
};

// Exhaustive histograms of codes A (each with 2^k words) in one batch, in the
// reported form of run_ancoding: counts[0] = 2^k, counts[i] = 2*raw[i]
// (u128 when wide, else doubled in 64 bits as super_a does).
cr_status batch_counts(cr_context *ctx, const std::vector<uintll_t> &As, uint32_t k, bool double_in_u64,
                       std::vector<CodeResult> &res, double &device_s) {
    std::vector<cr_an_code> codes(As.size());
    uint32_t stride = 1;
    for (size_t i = 0; i < As.size(); ++i) {
        uintll_t A = As[i];
        if ((A & (A - 1)) == 0) A = 1;  // power of two behaves as A = 1 (an_coding.cu:183-184)
        codes[i].A = A;
        codes[i].num_words = 1ull << k;
        stride = std::max(stride, pair_max_distance(A, 1ull << k) + 1);
    }
    std::vector<uint64_t> raw((size_t)As.size() * stride);
    double kms = 0;
    auto t0 = std::chrono::steady_clock::now();
    CR_TRY(cr_an_batch(ctx, codes.data(), (uint32_t)codes.size(), raw.data(), stride, nullptr, &kms));
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    device_s = kms * 1e-3;
    long double tot_pairs = 0;
    for (auto &c : codes) tot_pairs += (long double)c.num_words * (c.num_words - 1) / 2;
    res.assign(As.size(), CodeResult{});
    for (size_t i = 0; i < As.size(); ++i) {
        CodeResult &r = res[i];
        r.counts[0] = 1ull << k;
        for (uint32_t b = 1; b < stride && b < COUNTS_MAX_WIDTH; ++b) {
            const uint64_t v = raw[(size_t)i * stride + b];
            r.counts[b] = double_in_u64 ? (uint128_t)(uint64_t)(v << 1) : ((uint128_t)v << 1);
        }
        const long double share = ((long double)codes[i].num_words * (codes[i].num_words - 1) / 2) / tot_pairs;
        r.kernel_s = (double)(device_s * share);
        r.total_s = (double)(wall * share);
    }
    return CR_OK;
}

cr_status grid_counts(cr_context *ctx, int gdim, uint32_t k, uintll_t A, uintll_t m1, uintll_t m2, uint_t cc,
                      CodeResult &r) {
    std::vector<uint64_t> hist(65, 0), rep(2 * (size_t)cc);
    double kms = 0;
    auto t0 = std::chrono::steady_clock::now();
    CR_TRY(cr_an_grid_ref(ctx, gdim, k, m1, m2, A, hist.data(), 65, &kms));
    CR_TRY(cr_an_grid_report(gdim, k, m1, m2, hist.data(), cc, rep.data()));
    r = CodeResult{};
    for (uint_t i = 0; i < cc && i < COUNTS_MAX_WIDTH; ++i)
        r.counts[i] = ((uint128_t)rep[2 * i + 1] << 64) | rep[2 * i];
    r.kernel_s = kms * 1e-3;
    r.total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CR_OK;
}

std::string u128s(uint128_t v) {
    std::stringstream ss;
    ss << v;
    return ss.str();
}

// ancoding_search_super_A of distance_distribution/src/main.cpp:131-177
cr_status search_dd(cr_context *ctx, const cr_supera_request &q, std::string &out, double &device_s) {
    if (q.h_bits < 2 || q.h_bits > 40) return fail(CR_ERR_INVALID_ARGUMENT, "h_bits=%u outside [2, 40]", q.h_bits);
    const uint32_t n = q.k;
    std::vector<uintll_t> As;
    for (uintll_t A = (1ull << (q.h_bits - 1)) + 1; A <= (1ull << q.h_bits) - 1; A += 2) As.push_back(A);
    std::vector<CodeResult> res;
    device_s = 0;
    if (n <= 16) {
        CR_TRY(batch_counts(ctx, As, n, false, res, device_s));
    } else {
        if (q.with_grid != 1 && q.with_grid != 2) return fail(CR_ERR_INVALID_ARGUMENT, "k > 16 needs with_grid 1 or 2");
        res.resize(As.size());
        for (size_t i = 0; i < As.size(); ++i) {
            CR_TRY(grid_counts(ctx, q.with_grid, n, As[i], q.m1, q.m2, n + ceil_h(As[i]) + 1, res[i]));
            device_s += res[i].kernel_s;
        }
    }
    std::stringstream ss;
    uint128_t mincb = 0, mxmincb = static_cast<uint128_t>(-1), mxmincb2 = static_cast<uint128_t>(-1);
    uintll_t minb = 0, mxminb = 0, mxminb2 = 0, A2 = 0, superA = 0;
    double ttimes = 0;
    for (size_t i = 0; i < As.size(); ++i) {
        const uintll_t A = As[i];
        const uintll_t Ae = (A & (A - 1)) == 0 ? 1 : A;
        const uint_t cc = n + ceil_h(Ae) + 1;
        lowest(res[i].counts, cc / 2, minb, mincb);
        if (mxminb < minb || (mxminb == minb && mxmincb > mincb)) {
            if (mxminb != minb) {
                mxminb2 = mxminb;
                mxmincb2 = mxmincb;
                A2 = superA;
            }
            mxminb = minb;
            mxmincb = mincb;
            superA = A;
            ss << A << ": c[" << minb << "] = " << u128s(mincb) << std::endl;
        }
        ttimes += res[i].kernel_s;  // times[1] of run_ancoding is the kernel time
    }
    ss << "n," << n << ",h," << q.h_bits << ",superA," << superA << ",prevA," << A2 << ",c[" << mxminb << "],"
       << u128s(mxmincb) << ",c[" << mxminb2 << "]," << u128s(mxmincb2) << ",time[s]," << ttimes
       << (n <= 16 ? ",exact" : ",grid") << std::endl;
    out = ss.str();
    return CR_OK;
}

// device line of super_a's header (super_a/src/main.cpp:115-137)
std::string device_header(int dev) {
    hipDeviceProp_t p;
    size_t f = 0, t = 0;
    int rt = 0;
    std::stringstream info;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "";
    (void)hipMemGetInfo(&f, &t);
    (void)hipRuntimeGetVersion(&rt);
    info << '"' << p.name << '"' << ", \"CC\", " << p.major << '.' << p.minor << ", \"PCI Bus ID\", " << p.pciBusID
         << ", \"PCI Device ID\", " << p.pciDeviceID << ", \"Multiprocessors\", " << p.multiProcessorCount
         << ", \"Memory [MiB]\", " << t / 1048576 << ", \"MemoryFree [MiB]\", " << f / 1048576
         << ", \"ECC enabled\", " << p.ECCEnabled << ", \"MemClock [MHz]\", " << p.memoryClockRate / 1000
         << ", \"GPUClock [MHz]\", " << p.clockRate / 1000 << ", \"HIP Runtime\", " << rt;
    return info.str();
}

// ancoding_search_super_A(_by_file) of distance_distribution_super_a/src/main.cpp:194-297
cr_status search_supera(cr_context *ctx, const cr_supera_request &q, std::string &out, double &device_s) {
    const uint32_t n = q.k;
    std::vector<uintll_t> As;
    if (q.nAs) {
        for (uint32_t i = 0; i < q.nAs; ++i)
            if (q.As[i]) As.push_back(q.As[i]);  // value 0 is ignored (super_a/src/main.cpp:184-185)
    } else {
        for (uintll_t A = q.A_start; A <= q.A_end; A += 2) As.push_back(A);
    }
    if (As.empty()) return fail(CR_ERR_INVALID_ARGUMENT, "no A to search");
    Flags flags;
    flags.n = n;
    flags.with_grid = q.with_grid;
    flags.mc_iterations = (int)q.m1;
    flags.mc_iterations_2 = (int)q.m2;
    flags.verbose = q.verbose;
    std::vector<CodeResult> res;
    device_s = 0;
    if (q.with_grid == 0) {
        CR_TRY(batch_counts(ctx, As, n, true, res, device_s));
    } else {
        res.resize(As.size());
        for (size_t i = 0; i < As.size(); ++i) {
            uintll_t A = As[i];
            const uint_t h = floor_h(A);
            if ((A & (A - 1)) == 0) A = 1;
            const uintll_t msgs = 1ull << n;
            const uintll_t m1 = q.m1 > msgs ? msgs : q.m1;
            const uintll_t m2 = q.with_grid == 1 ? msgs : (q.m2 > msgs ? msgs : q.m2);
            CR_TRY(grid_counts(ctx, q.with_grid, n, A, m1, m2, n + h + 1, res[i]));
            device_s += res[i].kernel_s;
        }
    }
    std::stringstream ss;
    ss << device_header(ctx->device) << "\n";
    const char sep = ',';
    uint128_t mincb = 0, mxmincb = static_cast<uint128_t>(-1), mxmincb2 = static_cast<uint128_t>(-1);
    uintll_t minb = 0, mxminb = 0, mxminb2 = 0, superA = 0, superA2 = 0;
    double ttimes = 0;
    unsigned id = 0;
    for (size_t i = 0; i < As.size(); ++i) {
        const uintll_t A = As[i];
        const uint_t h = floor_h(A);
        lowest(res[i].counts, (n + h + 1 + 1) / 2, minb, mincb);
        if (mxminb < minb || (mxminb == minb && mxmincb > mincb)) {
            mxminb = minb;
            mxmincb = mincb;
            superA = A;
        }
        if (mxminb2 < minb || (mxminb2 == minb && mxmincb2 > mincb)) {
            mxminb2 = minb;
            mxmincb2 = mincb;
            superA2 = A;
        }
        double times[2] = {res[i].kernel_s, res[i].total_s};
        ttimes += times[1];
        supera_row(ss, A, h, flags, times, res[i].counts, &id);
        ss << sep << mxminb << sep << u128s(mxmincb) << sep << superA << sep << mxminb2 << sep << u128s(mxmincb2)
           << sep << superA2;
        ss << "\n";
        if (((A + 1) & ~A) == (A + 1)) {  // A+1 a power of two: next A starts a new bit-width class
            mxmincb2 = static_cast<uint128_t>(-1);
            mxminb2 = 0;
            superA2 = 0;
        }
    }
    std::stringstream file;
    file << "\"Total Time [s]\"," << ttimes << "\n" << ss.str();
    out = file.str();
    return CR_OK;
}

}  // namespace

extern "C" cr_status cr_supera_search(cr_context *ctx, const cr_supera_request *req, char **text, double *device_s) {
    clear_error();
    if (!ctx || !req || !text) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    *text = nullptr;
    if (req->k < 1 || req->k > 40) return fail(CR_ERR_UNSUPPORTED, "k=%u outside [1, 40]", req->k);
    std::string out;
    double dev_s = 0;
    cr_status s;
    if (req->style == 0)
        s = search_dd(ctx, *req, out, dev_s);
    else if (req->style == 1)
        s = search_supera(ctx, *req, out, dev_s);
    else
        return fail(CR_ERR_INVALID_ARGUMENT, "unknown style %u", req->style);
    if (s != CR_OK) return s;
    *text = (char *)malloc(out.size() + 1);
    if (!*text) return fail(CR_ERR_OUT_OF_MEMORY, "host allocation failed");
    memcpy(*text, out.c_str(), out.size() + 1);
    if (device_s) *device_s = dev_s;
    return CR_OK;
}

extern "C" void cr_free(void *p) { free(p); }
