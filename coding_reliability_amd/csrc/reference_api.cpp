// The reference's C++ entry points (algorithms.h of distance_distribution and
// distance_distribution_super_a) on the MI355X engine, with the reference's
// result post-processing and output formats (globals.cpp of both programs).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "coding_reliability.h"
#include "cr_reference_api.h"
#include "reference_tables.h"

using namespace std;

int Statistics::add(const std::string &label, const std::string &unit) {
    rows_.push_back(Row{label, unit, 0});
    return (int)rows_.size() - 1;
}

void Statistics::set(int i, double value) { rows_.at(i).value = value; }

void Statistics::check_index(int i) const {
    if (i < 0 || i >= getLength()) {
        fprintf(stderr, "Index out of range (%d/%d).\n", i, getLength());
        throw std::out_of_range("");
    }
}

namespace {

const int kOffset = 2;  // first bin of the error columns (globals.cpp:24)

std::mutex g_mu;
cr_context *g_ctx[64] = {};
std::recursive_mutex g_dev_mu[64];  // one call at a time per device context

[[noreturn]] void raise(const char *what) {
    throw std::runtime_error(std::string(what) + ": " + cr_last_error());
}

void check(cr_status s, const char *what) {
    if (s != CR_OK) raise(what);
}

// One context per device for the whole process (the reference created and
// reset the device on every call).  A lease holds the device's call lock for
// the duration of an entry point: the reference functions may be called from
// several threads, and a context serves one call at a time.
cr_context *context_for(int dev) {
    if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx[dev]) check(cr_context_create(dev, &g_ctx[dev]), "cr_context_create");
    return g_ctx[dev];
}

std::recursive_mutex &device_mutex(int dev) {
    if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
    return g_dev_mu[dev];
}

struct Lease {
    std::unique_lock<std::recursive_mutex> lock;
    cr_context *ctx;
    explicit Lease(int dev) : lock(device_mutex(dev)), ctx(context_for(dev)) {}
};

struct WallClock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// "Device" line appended to the exhaustive/grid reports (an_coding.cu:212-215).
std::string device_line(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "";
    std::stringstream ss;
    ss << "\"Device\", " << dev << ", \"MemClock [MHz]\", " << p.memoryClockRate / 1000 << ", \"GPUClock [MHz]\", "
       << p.clockRate / 1000 << std::endl;
    return ss.str();
}

template <typename T>
T binom(T n, T k) {  // globals.h:49-66
    T res = 1;
    if (k > n - k) k = n - k;
    for (T i = 0; i < k; ++i) {
        res *= (n - i);
        res /= (i + 1);
    }
    return res;
}

uint_t ceil_log2(uintll_t A) { return (uint_t)ceil(log((double)A) / log(2.0)); }

// Table of the reference's golden AN vectors (solutions.h:26-114).
const unsigned long long *an_solution(uintll_t A, uint_t n, unsigned *len) {
    using namespace cr::tables;
    if (A == 61) {
        if (n == 8) return *len = solution_an8_A61_len, solution_an8_A61;
        if (n == 16) return *len = solution_an16_A61_len, solution_an16_A61;
        if (n == 24) return *len = solution_an24_A61_len, solution_an24_A61;
    } else if (A == 641) {
        if (n == 8) return *len = solution_an8_A641_len, solution_an8_A641;
        if (n == 16) return *len = solution_an16_A641_len, solution_an16_A641;
        if (n == 24) return *len = solution_an24_A641_len, solution_an24_A641;
    }
    return nullptr;
}

// get_sol_hamming (globals.cpp:30-63): reported-form Hamming solution.
void sol_hamming(uint128_t *sol, uint_t n, uint_t *h_out) {
    using namespace cr::tables;
    *h_out = n == 8 ? 5 : (n == 16 || n == 24) ? 6 : 7;
    const uint_t h = *h_out;
    const unsigned long long *base;
    switch (n) {
    case 8: base = sol_hamming_n8; break;
    case 16: base = sol_hamming_n16; break;
    case 24: base = sol_hamming_n24; break;
    case 32: base = sol_hamming_n32; break;
    default: throw std::runtime_error("Wrong n for hamming.");
    }
    sol[0] = 1ull << n;
    sol[1] = (n + h) * sol[0];
    for (uint_t i = 3; i < n + h + 1; i += 2) {
        if (i + 1 < n + h + 1) {
            sol[i + 1] = (uint128_t)base[i + 1] << n;
            sol[i] = uint128_t(i + 1) * sol[i + 1] + uint128_t(n + h - i + 1) * sol[i - 1];
        } else {
            sol[i] = uint128_t(n + h - i + 1) * sol[i - 1];
        }
    }
}

// process_result (globals.cpp:185-244): the table written to stdout and, with
// a prefix, to <prefix>_n<n>_h<h>.csv.  Returns the histogram total.
long double write_result(uint128_t *counts, const Statistics &stats, uint_t n, uint_t h, const char *file_prefix,
                         double *errors_abs, double *errors_rel, const std::string &app) {
    std::stringstream ss;
    const char sep = ',';
    long double total = 0;
    for (uint_t i = 0; i < n + h + 1; ++i) total += static_cast<long double>(counts[i]);
    long double base1 = pow(2.0, n);
    double mxabs = 0, mxrel = 0;
    for (uint_t i = 0; i < n + h + 1; ++i) {
        const long double base2 = binom(static_cast<long double>(n + h), static_cast<long double>(i));
        const long double base = base1 * base2;
        const long double prob = counts[i] / base;
        if (n >= 32 && counts[i] > (uint128_t(1) << 64))
            ss << setw(4) << i << sep << setw(12) << setprecision(12) << static_cast<long double>(counts[i]);
        else
            ss << setw(4) << i << sep << setw(14) << static_cast<uintll_t>(counts[i]);
        ss << sep << setw(14) << prob << sep << setw(14) << base;
        if (errors_abs) {
            ss << sep << setw(13) << setprecision(7) << errors_abs[i];
            if (errors_abs[i] > mxabs) mxabs = errors_abs[i];
        }
        if (errors_rel) {
            ss << sep << setw(13) << setprecision(7) << errors_rel[i];
            if (errors_rel[i] > mxrel) mxrel = errors_rel[i];
        }
        ss << endl;
    }
    ss << endl << endl;
    for (int i = 0; i < stats.getLength(); ++i)
        ss << i << sep << '"' << stats.getLabel(i) << '"' << sep << stats.getAverage(i) << sep << stats.getUnit(i)
           << endl;
    ss << endl << "\"Total\"" << sep << total << endl;
    if (errors_abs) ss << "\"MaxAbsErr\"" << sep << mxabs << endl;
    if (errors_rel) ss << "\"MaxRelErr\"" << sep << mxrel << endl;
    ss << endl;
    ss << app;
    cout << ss.str();
    if (file_prefix) {
        char fname[256];
        snprintf(fname, sizeof(fname), "%s_n%u_h%u.csv", file_prefix, n, h);
        ofstream fp(fname);
        fp << ss.str();
        fp.close();
        cout << "File " << fname << " written." << endl << endl;
    }
    return total;
}

void printbits(uintll_t v, uintll_t n) {  // globals.cpp:306-315
    uintll_t mask = (1ull << n) - 1;
    uintll_t p = n;
    v &= mask;
    for (; v & mask; v <<= 1, --p) putchar('0' + ((v & mask) >> (n - 1)));
    for (; p; --p) putchar('0');
}

Statistics make_stats(double kernel_ms, double total_ms, bool super_a_order) {
    // the drivers add "Kernel Runtime" first, then "Total Runtime"; values in ms, printed in s
    Statistics s;
    const int ik = s.add("Kernel Runtime", "s");
    const int it = s.add("Total Runtime", "s");
    s.set(ik, kernel_ms);
    s.set(it, total_ms);
    s.setFactorAll(0.001);
    (void)super_a_order;
    return s;
}

void min_bin(const uint128_t *counts, uint_t limit, uintll_t *minb, uint128_t *mincb) {
    if (!minb || !mincb) return;
    *minb = 0xFFFF;
    *mincb = static_cast<uint128_t>(-1);
    for (uint_t i = 1; i < limit; ++i)
        if (counts[i] != 0 && counts[i] < *mincb) {
            *minb = i;
            *mincb = counts[i];
            break;
        }
}

int device_of(int nr_dev) {
    (void)nr_dev;  // one GPU per process: the engine runs on device 0
    return 0;
}

// Shard size of the reference Hamming grid / MC kernels (Hamming::traits::Shards, hamming.h:74-82).
uintll_t hamming_shards(uintll_t n) {
    switch (n) {
    case 4:
    case 8: return 1;
    case 16: return 16;
    case 24: return 128;
    case 32: return 256;
    case 40: return 512;
    case 48: return 1024;
    default: throw std::runtime_error("Unsupported n.");
    }
}

// Samples run_ancoding_mc has already drawn: the raw histogram of samples
// [0, m) for (A, n) on one device with the fixed seed.  Sample j of word i
// depends only on the seed, i and j, so a later call with the same A and n
// and at least m samples draws samples [m, iterations) alone and adds them,
// with the histogram a fresh call would give.  ancoding_mc_search
// (main.cpp:179-202) multiplies the iterations by 3 per step: every sample is
// drawn once instead of once per later step.
struct McMemo {
    uint64_t A = 0, n = 0, m = 0;
    std::vector<uint64_t> hist;
};
std::mutex g_mc_mu;
McMemo g_mc_memo;  // device 0, the one device of the process

// run_hamming_grid / run_hamming_mc output h (hamming_grid.cu:127, hamming_mc.cu:133)
uint_t hamming_h_ref(uintll_t n) { return n == 8 ? 5 : (n < 32 ? 6 : 7); }

// Merge/sphere post-processing of the sampled Hamming drivers
// (hamming_grid.cu:182-202, hamming_mc.cu:194-214).
void hamming_sampled_report(const std::vector<uint64_t> &h, uintll_t N, long double factor, int with_1bit,
                            uint128_t *counts) {
    const uint_t cc = (uint_t)N + 1;
    counts[0] = static_cast<uint128_t>(factor * h[0]);
    for (uint_t i = 2; i < cc; i += 2) counts[i] = static_cast<uint128_t>(factor * (h[i] + h[i - 1]));
    if (with_1bit)
        for (uint_t i = 1; i < cc; i += 2)
            counts[i] = i + 1 < cc ? uint128_t(i + 1) * counts[i + 1] + uint128_t(N - i + 1) * counts[i - 1]
                                   : uint128_t(N - i + 1) * counts[i - 1];
}

}  // namespace

// ---------------------------------------------------------------------------
// error helpers and result writers (distance_distribution/src/globals.cpp)
// ---------------------------------------------------------------------------

double get_rel_error_hamming(uint_t n, uint128_t *tgt, int offset, int w1bit, double *errors) {
    uint128_t sol[64] = {0};
    double max_err = 0, err = 0, p;
    uint_t h = 0;
    sol_hamming(sol, n, &h);
    if (errors) memset(errors, 0, (n + h + 1) * sizeof(double));
    if (w1bit == 0 && (offset & 1) == 1) offset += 1;
    for (uint_t k = offset; k < n + h + 1; k += w1bit ? 1 : 2) {
        p = static_cast<double>(sol[k]);
        if (p > 0.)
            err = fabs(static_cast<double>(tgt[k]) / p - 1.0);
        else if (tgt[k] > 0)
            err = 1.0;
        else
            err = 0.0;
        if (errors) errors[k] = err;
        if (err > max_err) max_err = err;
    }
    return max_err;
}

double get_abs_error_hamming(uint_t n, uint128_t *tgt, int offset, int w1bit, double *errors) {
    uint128_t sol[64] = {0};
    double max_err = 0, err = 0, p;
    uint_t h = 0;
    sol_hamming(sol, n, &h);
    if (errors) memset(errors, 0, (n + h + 1) * sizeof(double));
    if (w1bit == 0 && (offset & 1) == 1) offset += 1;
    for (uint_t k = offset; k < n + h + 1; k += w1bit ? 1 : 2) {
        p = static_cast<double>(sol[k] > tgt[k] ? sol[k] - tgt[k] : tgt[k] - sol[k]);
        err = p / (pow(2.0, n) * binom(static_cast<double>(n) + h, static_cast<double>(k)));
        if (errors) errors[k] = err;
        if (err > max_err) max_err = err;
    }
    return max_err;
}

double get_rel_error_AN(uintll_t A, uint_t n, uint128_t *tgt, int offset, double *errors) {
    unsigned len = 0;
    const unsigned long long *sol = an_solution(A, n, &len);
    if (!sol) return -1.0;
    const uint_t h = A == 641 ? 10 : 6;
    double max_err = 0, err = 0, p;
    if (errors) memset(errors, 0, (n + h + 1) * sizeof(double));
    for (uint_t k = offset; k < n + h && k < len; ++k) {
        p = static_cast<double>(sol[k]);
        if (p > 0.)
            err = fabs(static_cast<double>(tgt[k]) / p - 1.0);
        else if (tgt[k] > 0)
            err = 1.0;
        else
            err = 0.0;
        if (errors) errors[k] = err;
        if (err > max_err) max_err = err;
    }
    return max_err;
}

double get_abs_error_AN(uintll_t A, uint_t n, uint128_t *tgt, int offset, double *errors) {
    unsigned len = 0;
    const unsigned long long *sol = an_solution(A, n, &len);
    if (!sol) return -1.0;
    const uint_t h = A == 641 ? 10 : 6;
    double max_err = 0, err = 0, p;
    if (errors) memset(errors, 0, (n + h + 1) * sizeof(double));
    for (uint_t k = offset; k < n + h && k < len; ++k) {
        const uint128_t s = sol[k];
        p = static_cast<double>(s > tgt[k] ? s - tgt[k] : tgt[k] - s);
        err = p / (pow(2.0, n) * binom(static_cast<double>(n) + h, static_cast<double>(k)));
        if (errors) errors[k] = err;
        if (err > max_err) max_err = err;
    }
    return max_err;
}

void process_result_hamming(uint128_t *counts, Statistics stats, uint_t n, uint_t h, const char *file_prefix) {
    long double total = write_result(counts, stats, n, h, file_prefix, nullptr, nullptr, "");
    cout << endl << "Sum counts = " << setw(12) << setprecision(12) << total << endl;
}

void process_result_hamming_mc(uint128_t *counts, Statistics stats, uint_t n, uint_t h, int with_1bit,
                               uint_t iterations, const char *file_prefix) {
    char tmp[192];
    double errors_abs[64] = {0}, errors_rel[64] = {0};
    get_abs_error_hamming(n, counts, kOffset, with_1bit, errors_abs);
    get_rel_error_hamming(n, counts, kOffset, with_1bit, errors_rel);
    long double total;
    if (file_prefix) {
        snprintf(tmp, sizeof(tmp), "%s_m%u", file_prefix, iterations);
        total = write_result(counts, stats, n, h, tmp, errors_abs, errors_rel, "");
    } else {
        total = write_result(counts, stats, n, h, nullptr, errors_abs, errors_rel, "");
    }
    cout << endl
         << "Sum counts = " << setw(12) << setprecision(12) << total
         << " (with estimation factor 2^n*counts[d]/iterations)" << endl;
}

void process_result_ancoding(uint128_t *counts, Statistics stats, uint_t n, uint_t A, const char *file_prefix,
                             std::string app) {
    char tmp[192];
    const uint_t h = ceil_log2(A);
    long double total;
    if (file_prefix) {
        snprintf(tmp, sizeof(tmp), "%s_A%u", file_prefix, A);
        total = write_result(counts, stats, n, h, tmp, nullptr, nullptr, app);
    } else {
        total = write_result(counts, stats, n, h, nullptr, nullptr, nullptr, app);
    }
    cout << endl << "Sum counts = " << setw(12) << setprecision(12) << total << endl;
    cout << " A = ";
    cout.flush();
    printbits(A, n);
    fflush(stdout);
    cout << " = " << A << endl;
}

void process_result_ancoding_mc(uint128_t *counts, Statistics stats, uint_t n, uint_t A, uint_t iterations,
                                const char *file_prefix, std::string app) {
    char tmp[192];
    double errors_abs[64] = {0}, errors_rel[64] = {0};
    get_abs_error_AN(A, n, counts, kOffset, errors_abs);
    get_rel_error_AN(A, n, counts, kOffset, errors_rel);
    const uint_t h = static_cast<uint_t>(ceil(log2(A)));
    long double total;
    if (file_prefix) {
        snprintf(tmp, sizeof(tmp), "%s_m%u_A%u", file_prefix, iterations, A);
        total = write_result(counts, stats, n, h, tmp, errors_abs, errors_rel, app);
    } else {
        total = write_result(counts, stats, n, h, nullptr, errors_abs, errors_rel, app);
    }
    cout << endl
         << "Sum counts = " << setw(12) << setprecision(12) << total
         << " (with estimation factor 2^n*counts[d]/iterations)" << endl;
    cout << " A = ";
    cout.flush();
    printbits(A, n);
    fflush(stdout);
    cout << " = " << A << endl;
}

std::ostream &operator<<(std::ostream &dest, uint128_t value) {  // globals.cpp:320-341
    std::ostream::sentry s(dest);
    if (s) {
        char buffer[128];
        char *d = buffer + sizeof(buffer);
        uint128_t tmp = value;
        do {
            *--d = "0123456789"[tmp % 10];
            tmp /= 10;
        } while (tmp != 0);
        const int len = (int)(buffer + sizeof(buffer) - d);
        if (dest.rdbuf()->sputn(d, len) != len) dest.setstate(std::ios_base::badbit);
    }
    return dest;
}

// ---------------------------------------------------------------------------
// distance_distribution/src/algorithms.h entry points
// ---------------------------------------------------------------------------

void run_ancoding(uintll_t n, uintll_t A, int verbose, double *times, uintll_t *minb, uint128_t *mincb,
                  int file_output, int nr_dev) {
    // an_coding.cu:181-327
    if ((A & (A - 1)) == 0) A = 1;
    const uint_t h = ceil_log2(A);
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    const std::string app = device_line(dev);
    WallClock wc;
    const uint_t count_counts = (uint_t)n + h + 1;
    std::vector<uint64_t> raw(std::max<uint_t>(count_counts, 65), 0);
    double kms = 0;
    check(cr_an_pairs(ctx, A, 1ull << n, raw.data(), (uint32_t)raw.size(), &kms), "run_ancoding");
    const double tms = wc.ms();
    uint128_t counts[64] = {0};
    counts[0] = 1ull << n;
    for (uint_t i = 1; i < count_counts && i < 64; ++i) counts[i] = static_cast<uint128_t>(raw[i]) << 1;
    min_bin(counts, count_counts / 2, minb, mincb);
    Statistics stats = make_stats(kms, tms, false);
    if (times) {
        times[0] = stats.getAverage(1);  // total
        times[1] = stats.getAverage(0);  // kernel
    }
    if (verbose || file_output) {
        const char *prefix = file_output ? (nr_dev == 4 ? "ancoding_4gpu" : "ancoding_gpu") : nullptr;
        process_result_ancoding(counts, stats, (uint_t)n, (uint_t)A, prefix, app);
    }
}

double run_ancoding_grid(int gdim, uintll_t n, uintll_t iterations, uintll_t iterations2, uintll_t A, int verbose,
                         double *times, uintll_t *minb, uint128_t *mincb, int file_output, int nr_dev) {
    // an_coding_grid.cu:215-362
    if ((A & (A - 1)) == 0) A = 1;
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    if (verbose > 1)
        printf("Start AN-Coding Algorithm - %dD Grid with %llu x %llu points\n", gdim, iterations,
               gdim == 2 ? iterations2 : 1);
    const std::string app = device_line(dev);
    WallClock wc;
    const uintll_t msgs = 1ull << n;
    iterations = iterations > msgs ? msgs : iterations;
    iterations2 = gdim == 1 ? msgs : (iterations2 > msgs ? msgs : iterations2);
    const uint_t count_counts = (uint_t)n + ceil_log2(A) + 1;
    std::vector<uint64_t> hist(65, 0);
    double kms = 0;
    check(cr_an_grid_ref(ctx, gdim, (uint32_t)n, iterations, iterations2, A, hist.data(), 65, &kms),
          "run_ancoding_grid");
    const double tms = wc.ms();
    std::vector<uint64_t> rep(2 * count_counts);
    check(cr_an_grid_report(gdim, (uint32_t)n, iterations, iterations2, hist.data(), count_counts, rep.data()),
          "run_ancoding_grid");
    uint128_t counts[64] = {0};
    for (uint_t i = 0; i < count_counts && i < 64; ++i) counts[i] = ((uint128_t)rep[2 * i + 1] << 64) | rep[2 * i];
    const double max_err = get_rel_error_AN(A, (uint_t)n, counts, 0);
    min_bin(counts, count_counts / 2, minb, mincb);
    Statistics stats = make_stats(kms, tms, false);
    if (verbose || file_output) {
        char fname[256];
        snprintf(fname, sizeof(fname), "ancoding_grid%dd_%s", gdim, nr_dev == 4 ? "4gpu" : "gpu");
        process_result_ancoding_mc(counts, stats, (uint_t)n, (uint_t)A, (uint_t)iterations,
                                   file_output ? fname : nullptr, app);
    }
    if (times) {
        times[0] = stats.getAverage(1);
        times[1] = stats.getAverage(0);
    }
    return max_err;
}

double run_ancoding_mc(uintll_t n, uintll_t iterations, uintll_t A, int verbose, double *times, uintll_t *minb,
                       uint128_t *mincb, int file_output, int nr_dev) {
    // an_coding_mc.cu:205-343 with Philox samples (seed 16384, the reference's device-0 seed)
    if ((A & (A - 1)) == 0) A = 1;
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    WallClock wc;
    const uintll_t msgs = 1ull << n;
    const uint_t count_counts = (uint_t)n + ceil_log2(A) + 1;
    iterations = iterations > msgs ? msgs : iterations;
    std::vector<uint64_t> hist(65, 0);
    double kms = 0;
    McMemo memo;
    {
        std::lock_guard<std::mutex> g(g_mc_mu);
        memo = g_mc_memo;
    }
    if (dev == 0 && memo.A == A && memo.n == n && memo.m <= iterations && memo.hist.size() == hist.size()) {
        // only the samples beyond the memo's
        const uint64_t bounds[2] = {memo.m, iterations};
        check(cr_an_mc_ranges(ctx, A, (uint32_t)n, bounds, 1, 16384, hist.data(), 65, &kms), "run_ancoding_mc");
        for (size_t i = 0; i < hist.size(); ++i) hist[i] += memo.hist[i];
    } else {
        check(cr_an_mc(ctx, A, (uint32_t)n, iterations, 16384, hist.data(), 65, &kms), "run_ancoding_mc");
    }
    if (dev == 0) {
        std::lock_guard<std::mutex> g(g_mc_mu);
        g_mc_memo.A = A;
        g_mc_memo.n = n;
        g_mc_memo.m = iterations;
        g_mc_memo.hist = hist;
    }
    const double tms = wc.ms();
    uint128_t counts[64] = {0};
    for (uint_t i = 0; i < count_counts && i < 64; ++i)
        counts[i] = static_cast<uint128_t>(static_cast<long double>(powl(2.0L, (long double)n) / iterations * hist[i]));
    const double max_err = get_rel_error_AN(A, (uint_t)n, counts, 0);
    min_bin(counts, count_counts / 2, minb, mincb);
    Statistics stats = make_stats(kms, tms, false);
    if (verbose || file_output) {
        const char *prefix = file_output ? (nr_dev == 4 ? "ancoding_mc_4gpu" : "ancoding_mc_gpu") : nullptr;
        process_result_ancoding_mc(counts, stats, (uint_t)n, (uint_t)A, (uint_t)iterations, prefix);
    }
    if (times) {
        times[0] = stats.getAverage(1);
        times[1] = stats.getAverage(0);
    }
    return max_err;
}

void run_hamming(uintll_t n, int with_1bit, int file_output, int nr_dev) {
    // hamming.cu:97-224: weights of all 2^n codewords, merge, spheres, << n
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    WallClock wc;
    const uint_t r = cr_hamming_parity_bits((uint32_t)n);
    const uint_t N = (uint_t)n + r + 1, h = r + 1;
    std::vector<uint64_t> w(N + 1, 0);
    cr_ham_request req{};
    req.k = (uint32_t)n;
    req.source = CR_HAM_RANGE;
    req.first = 0;
    req.count = 1ull << n;
    double kms = 0;
    check(cr_hamming_weights(ctx, &req, w.data(), N + 1, &kms), "run_hamming");
    std::vector<uint64_t> rep(2 * (N + 1));
    check(cr_hamming_report((uint32_t)n, w.data(), with_1bit, rep.data()), "run_hamming");
    const double tms = wc.ms();
    uint128_t counts[130] = {0};
    for (uint_t i = 0; i <= N; ++i) counts[i] = ((uint128_t)rep[2 * i + 1] << 64) | rep[2 * i];
    Statistics stats = make_stats(kms, tms, false);
    const char *prefix = file_output ? (nr_dev == 4 ? "hamming_4gpu" : "hamming_gpu") : nullptr;
    process_result_hamming(counts, stats, (uint_t)n, h, prefix);
}

double run_hamming_grid(uintll_t n, int with_1bit, uintll_t iterations, int file_output, int nr_dev) {
    // hamming_grid.cu:91-219: sample points x = shard*S*(u64)step + k*step
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    hipDeviceProp_t p;
    const char *name = hipGetDeviceProperties(&p, dev) == hipSuccess ? p.name : "";
    printf("Start 1D Hamming-Coding Algorithm - 1D Grid with %llu iterations\n", iterations);
    printf("Found %d CUDA devices (%s).\n", 1, name);
    WallClock wc;
    const uintll_t msgs = 1ull << n;
    const uintll_t S = hamming_shards(n);
    iterations = iterations > msgs ? msgs : iterations;
    const uintll_t shards = iterations / S;
    const double step = (double)(1.0L * msgs / iterations);  // long double argument narrowed to the kernel's double
    std::vector<uint64_t> xs;
    xs.reserve(shards * S);
    for (uintll_t sh = 0; sh < shards; ++sh) {
        const uintll_t w = sh * S * (uintll_t)step;
        for (uintll_t k = 0; k < S; ++k) {
            uintll_t x = w;
            x += k * step;
            xs.push_back(x);
        }
    }
    const uint_t h = hamming_h_ref(n);
    const uintll_t N = n + h;
    std::vector<uint64_t> w(std::max<size_t>(N + 2, 130), 0);
    cr_ham_request req{};
    req.k = (uint32_t)n;
    req.source = CR_HAM_LIST;
    req.count = xs.size();
    req.list = xs.data();
    double kms = 0;
    check(cr_hamming_weights(ctx, &req, w.data(), (uint32_t)w.size(), &kms), "run_hamming_grid");
    const double tms = wc.ms();
    uint128_t counts[130] = {0};
    hamming_sampled_report(w, N, powl(2.0L, (long double)n), with_1bit, counts);
    const double max_abs = get_abs_error_hamming((uint_t)n, counts, 0, with_1bit, nullptr);
    Statistics stats = make_stats(kms, tms, false);
    process_result_hamming_mc(counts, stats, (uint_t)n, h, with_1bit, (uint_t)iterations,
                              file_output ? "hamming_mc" : nullptr);
    return max_abs;
}

double run_hamming_mc(uintll_t n, int with_1bit, uintll_t iterations, int file_output, int nr_dev) {
    // hamming_mc.cu:98-231 with Philox words (seed 1337, the reference's device-0 seed)
    const int dev = device_of(nr_dev);
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    hipDeviceProp_t p;
    const char *name = hipGetDeviceProperties(&p, dev) == hipSuccess ? p.name : "";
    printf("Start Hamming-Coding Algorithm - Monte Carlo with %llu iterations\n", iterations);
    printf("Found %d CUDA devices (%s).\n", 1, name);
    WallClock wc;
    const uintll_t msgs = 1ull << n;
    const uintll_t S = hamming_shards(n);
    iterations = iterations > msgs ? msgs : iterations;
    const uintll_t samples = iterations / S * S;
    const uint_t h = hamming_h_ref(n);
    const uintll_t N = n + h;
    std::vector<uint64_t> w(std::max<size_t>(N + 2, 130), 0);
    cr_ham_request req{};
    req.k = (uint32_t)n;
    req.source = CR_HAM_MC;
    req.first = 0;
    req.count = samples;
    req.seed = 1337;
    req.stream = 0;
    double kms = 0;
    check(cr_hamming_weights(ctx, &req, w.data(), (uint32_t)w.size(), &kms), "run_hamming_mc");
    const double tms = wc.ms();
    uint128_t counts[130] = {0};
    hamming_sampled_report(w, N, powl(2.0L, (long double)(2 * n)) / iterations, with_1bit, counts);
    const double max_abs = get_abs_error_hamming((uint_t)n, counts, 0, with_1bit, nullptr);
    Statistics stats = make_stats(kms, tms, false);
    process_result_hamming_mc(counts, stats, (uint_t)n, h, with_1bit, (uint_t)iterations,
                              file_output ? "hamming_mc" : nullptr);
    return max_abs;
}

// ---------------------------------------------------------------------------
// Entry points of the reference's CPU implementations, computed on the GPU
// with those paths' own conventions.
// ---------------------------------------------------------------------------

namespace {

// Timing table of the CPU drivers: one row, "Total Runtime" (an_coding.cpp:106-109,
// hamming.cpp:96-99, hamming_native*.cpp).
Statistics cpu_stats(double total_ms) {
    Statistics s;
    const int it = s.add("Total Runtime", "s");
    s.set(it, total_ms);
    s.setFactorAll(0.001);
    return s;
}

// Reported-form Extended Hamming distribution of hamming_cpu<n> (hamming.cpp:43-93)
// for n in {8, 16, 24, 32, 40}: the non-extended weights of all 2^n data
// words, then the CPU driver's own post-processing, which differs from the
// GPU driver's (hamming.cu:183-208) where the code has words of weight 1 or
// 2: bins below 2 are not summed, bin 2 keeps the weight-2 count unshifted
// (only bins >= 3 are shifted by n), odd weights merge into the next even bin
// from bin 4 on, and the sphere of bin 3 uses bin 2.  apply_w<N>
// (hamming.cpp:25-41) takes 4, 5 or 6 of the 32-bit parity masks; its masks
// see only the low 32 data bits, so for n = 40 the weight of a word is that
// of its low 32 bits under the (39,32) code plus the popcount of its high 8
// bits, and the distribution is the (39,32) one convolved with C(8, j).
void hamming_cpu_counts(cr_context *ctx, uint_t n, int with_1bit, uint128_t *counts, uint_t count_counts) {
    const uint_t k = n == 40 ? 32 : n;
    const uint_t N = k + cr_hamming_parity_bits(k) + 1;
    std::vector<uint64_t> w(N + 1, 0);
    cr_ham_request req{};
    req.k = k;
    req.source = CR_HAM_RANGE;
    req.first = 0;
    req.count = 1ull << k;
    check(cr_hamming_weights(ctx, &req, w.data(), N + 1, nullptr), "run_hamming_cpu");
    if (n == 40) {
        std::vector<uint64_t> w40(N + 8 + 1, 0);
        for (uint_t a = 0; a <= N; ++a)
            for (uint_t j = 0; j <= 8; ++j) w40[a + j] += w[a] * binom<uint64_t>(8, j);
        w.swap(w40);
    }
    const uint_t h = n == 8 ? 5 : (n < 32 ? 6 : 7);
    const uint_t len = n + h;  // codeword bits incl. the overall parity bit
    const uint_t cc = std::min<uint_t>(count_counts, len + 1);
    std::vector<uint128_t> c(len + 1, 0);
    for (uint_t d = 2; d < cc && d < w.size(); ++d) c[d] = w[d];
    c[0] = (uint128_t)1 << n;
    c[1] = with_1bit ? (uint128_t)len * c[0] : 0;
    for (uint_t d = 4; d < cc; d += 2) {
        c[d] += c[d - 1];
        c[d - 1] = 0;
    }
    if (with_1bit)
        for (uint_t d = 3; d < cc; d += 2)
            c[d] = d < len ? (uint128_t)(len - d + 1) * c[d - 1] + (uint128_t)(d + 1) * c[d + 1]
                           : (uint128_t)(len - d + 1) * c[d - 1];
    for (uint_t d = 3; d < cc; ++d) c[d] <<= n;
    for (uint_t d = 0; d < cc; ++d) counts[d] = c[d];
}

// The brute-force drivers (hamming_native.cpp, hamming_native_short.cpp,
// hamming_native_short_ordering.cpp) instantiate their templates for 8, 16
// and, for every other n, 24 data bits, and fill the CNT_COUNTS = 8+13+1,
// 16+6+1 or 24+6+1 bins of that code.  By linearity their ordered-pair and
// sphere counts equal hamming_cpu's reported form (every codeword sees the
// same distance distribution); the ordering variant counts pairs a < b
// twice, so the diagonal's bins 0 and 1 stay zero.  Returns the code width
// used; throws where the reference would write past its n+h+1 bins.
uint_t hamming_native_counts(cr_context *ctx, uint_t n, int with_1bit, bool ordering, uint128_t *counts,
                             uint_t count_counts) {
    const uint_t code = (n == 8 || n == 16) ? n : 24;
    const uint_t cc = code + (code == 8 ? 5 : 6) + 1;
    if (cc > count_counts)
        throw std::runtime_error("run_hamming_cpu_native*: n=" + std::to_string(n) +
                                 " selects the 24-bit code, whose 31 bins exceed the n+h+1 = " +
                                 std::to_string(count_counts) + " the reference allocates");
    hamming_cpu_counts(ctx, code, with_1bit, counts, cc);
    if (ordering) counts[0] = counts[1] = 0;
    return code;
}

}  // namespace

void run_ancoding_cpu(uintll_t n, uintll_t A, int verbose, double *times, uintll_t *minb, uint128_t *mincb,
                      int file_output) {
    // an_coding.cpp:104-142: count_counts = n + ceil(log2 A) + 1 with A as
    // given (no power-of-two mapping), counts[0] = 2^n (the diagonal,
    // an_coding.cpp:69) and counts[b] = 2 * #{i < j : d = b}.  For n < 6 the
    // reference's 64-word shard walks past 2^n and writes beyond its bins; here
    // the 2^n words of the code are used.
    if (A == 0) throw std::invalid_argument("run_ancoding_cpu: A must be >= 1");
    const uintll_t count_counts = n + (uintll_t)ceil(log((double)A) / log(2.0)) + 1;
    const int dev = 0;
    Lease lease(dev);
    WallClock wc;
    std::vector<uint64_t> raw(std::max<uintll_t>(count_counts, 65), 0);
    check(cr_an_pairs(lease.ctx, A, 1ull << n, raw.data(), (uint32_t)raw.size(), nullptr), "run_ancoding_cpu");
    std::vector<uint128_t> counts(count_counts, 0);
    counts[0] = 1ull << n;
    for (uintll_t i = 1; i < count_counts; ++i) counts[i] = static_cast<uint128_t>(raw[i]) << 1;
    Statistics stats = cpu_stats(wc.ms());
    if (verbose || file_output)
        process_result_ancoding(counts.data(), stats, (uint_t)n, (uint_t)A, file_output ? "ancoding_cpu" : nullptr);
    if (minb != nullptr && mincb != nullptr) {
        // an_coding.cpp:125-137: the CPU driver starts from mincb = 0xFFFFFFFF
        *minb = 0xFFFF;
        *mincb = 0xFFFFFFFF;
        for (uint_t i = 1; i < count_counts / 2; ++i)
            if (counts[i] != 0 && counts[i] < *mincb) {
                *minb = i;
                *mincb = counts[i];
                break;
            }
    }
    if (times != NULL) {
        // an_coding.cpp:138-142: the table has one row, so the second read
        // throws std::out_of_range as in the reference
        times[0] = stats.getAverage(0);
        times[1] = stats.getAverage(1);
    }
}

void run_hamming_cpu(uintll_t n, int with_1bit, int file_output) {
    // hamming.cpp:94-126
    const uintll_t h = n == 8 ? 5 : (n < 32 ? 6 : 7);
    std::vector<uint128_t> counts(n + h + 1, 0);
    printf("Start Hamming Coding Algorithm (CPU)\n");
    Lease lease(0);
    WallClock wc;
    if (n == 8 || n == 16 || n == 24 || n == 32 || n == 40)
        hamming_cpu_counts(lease.ctx, (uint_t)n, with_1bit, counts.data(), (uint_t)(n + h + 1));
    Statistics stats = cpu_stats(wc.ms());
    process_result_hamming(counts.data(), stats, (uint_t)n, (uint_t)h, file_output ? "hamming_cpu" : nullptr);
}

void run_hamming_cpu_native_short(uintll_t n, int with_1bit, int file_output) {
    // hamming_native_short.cpp:86-122
    const uintll_t h = n == 8 ? 5 : (n < 32 ? 6 : 7);
    std::vector<uint128_t> counts(n + h + 1, 0);
    printf("Start Hamming Coding Algorithm - Native Short Approach (CPU)\n");
    Lease lease(0);
    WallClock wc;
    hamming_native_counts(lease.ctx, (uint_t)n, with_1bit, false, counts.data(), (uint_t)(n + h + 1));
    Statistics stats = cpu_stats(wc.ms());
    process_result_hamming(counts.data(), stats, (uint_t)n, (uint_t)h,
                           file_output ? "hamming_cpu_native_short" : nullptr);
}

void run_hamming_cpu_native_short_ordering(uintll_t n, int with_1bit, int file_output) {
    // hamming_native_short_ordering.cpp:85-118
    const uintll_t h = n == 8 ? 5 : (n < 32 ? 6 : 7);
    std::vector<uint128_t> counts(n + h + 1, 0);
    printf("Start Hamming Coding Algorithm - Native Short Approach with Ordering (CPU)\n");
    Lease lease(0);
    WallClock wc;
    hamming_native_counts(lease.ctx, (uint_t)n, with_1bit, true, counts.data(), (uint_t)(n + h + 1));
    Statistics stats = cpu_stats(wc.ms());
    process_result_hamming(counts.data(), stats, (uint_t)n, (uint_t)h,
                           file_output ? "hamming_cpu_native_short_ordering" : nullptr);
}

void run_hamming_cpu_native(uintll_t n, int with_1bit, int file_output) {
    // hamming_native.cpp:130-155; with spheres the reference also prints its
    // "#Distances" table (hamming_native.cpp:109-118)
    const uintll_t h = n == 8 ? 5 : (n < 32 ? 6 : 7);
    std::vector<uint128_t> counts(n + h + 1, 0);
    printf("Start Hamming Coding Algorithm - Native Approach (CPU)\n");
    Lease lease(0);
    WallClock wc;
    const uint_t code = hamming_native_counts(lease.ctx, (uint_t)n, with_1bit, false, counts.data(),
                                              (uint_t)(n + h + 1));
    if (with_1bit) {
        const uint_t cc = code + (code == 8 ? 5 : 6) + 1;
        cout << dec << "\n#Distances:\n";
        for (uint_t i = 0; i < cc; ++i)
            cout << "  " << right << setw(2) << i << ": " << setw(13) << static_cast<uintll_t>(counts[i]) << '\n';
    }
    Statistics stats = cpu_stats(wc.ms());
    process_result_hamming(counts.data(), stats, (uint_t)n, (uint_t)h, file_output ? "hamming_cpu_native" : nullptr);
}

static void not_provided(const char *fn) {
    throw std::runtime_error(std::string(fn) + ": the cuRAND dumps are not part of this GPU library");
}

void test_curand(uintll_t, uintll_t, int) { not_provided("test_curand"); }
void test_curand_1d(uintll_t, uintll_t, int) { not_provided("test_curand_1d"); }
void test_curand_raw(uintll_t, uintll_t, int) { not_provided("test_curand_raw"); }

// ---------------------------------------------------------------------------
// distance_distribution_super_a entry points
// ---------------------------------------------------------------------------

void run_ancoding(uintll_t A, uint_t h, Flags &flags, double *times, uint128_t *counts) {
    // super_a/src/an_coding.cu:135-244: counts returned to the caller, doubled in 64 bits
    if ((A & (A - 1)) == 0) A = 1;
    const int dev = flags.dev >= 0 ? flags.dev : 0;
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    WallClock wc;
    const uint_t n = (uint_t)flags.n;
    const uint_t count_counts = n + h + 1;
    std::vector<uint64_t> raw(std::max<uint_t>(count_counts, 65), 0);
    double kms = 0;
    check(cr_an_pairs(ctx, A, 1ull << n, raw.data(), (uint32_t)raw.size(), &kms), "run_ancoding");
    const double tms = wc.ms();
    counts[0] = 1ull << n;
    for (uint_t i = 1; i < count_counts && i < COUNTS_MAX_WIDTH; ++i) counts[i] = raw[i] << 1;
    if (times) {
        times[0] = kms * 0.001;  // kernel
        times[1] = tms * 0.001;  // total
    }
}

void run_ancoding_grid(uintll_t A, uint_t h, Flags &flags, double *times, uint128_t *counts) {
    // super_a/src/an_coding_grid.cu:179-289
    if ((A & (A - 1)) == 0) A = 1;
    const int dev = flags.dev >= 0 ? flags.dev : 0;
    Lease lease(dev);
    cr_context *ctx = lease.ctx;
    WallClock wc;
    const uint_t n = (uint_t)flags.n;
    const uint_t count_counts = n + h + 1;
    const uintll_t msgs = 1ull << n;
    const uintll_t it1 = (uintll_t)flags.mc_iterations > msgs ? msgs : (uintll_t)flags.mc_iterations;
    const uintll_t it2 = flags.with_grid == 1 ? msgs
                                              : ((uintll_t)flags.mc_iterations_2 > msgs ? msgs
                                                                                        : (uintll_t)flags.mc_iterations_2);
    std::vector<uint64_t> hist(65, 0);
    double kms = 0;
    check(cr_an_grid_ref(ctx, flags.with_grid, n, it1, it2, A, hist.data(), 65, &kms), "run_ancoding_grid");
    const double tms = wc.ms();
    std::vector<uint64_t> rep(2 * std::max<uint_t>(count_counts, 1));
    check(cr_an_grid_report(flags.with_grid, n, it1, it2, hist.data(), count_counts, rep.data()), "run_ancoding_grid");
    for (uint_t i = 0; i < count_counts && i < COUNTS_MAX_WIDTH; ++i)
        counts[i] = ((uint128_t)rep[2 * i + 1] << 64) | rep[2 * i];
    if (times) {
        times[0] = kms * 0.001;
        times[1] = tms * 0.001;
    }
}

static std::string printbits_str(uintll_t v, uintll_t n) {  // super_a/src/globals.cpp:26-37
    std::string s;
    uintll_t mask = (1ull << n) - 1;
    uintll_t p = n;
    v &= mask;
    for (; v & mask; v <<= 1, --p) s += char('0' + ((v & mask) >> (n - 1)));
    for (; p; --p) s += '0';
    return s;
}

namespace cr {

// One CSV row of super_a's process_result with an explicit row counter.
void supera_row(std::stringstream &ss, uintll_t A, uint_t h, const Flags &flags, double *times, uint128_t *counts,
                unsigned *id) {
    // super_a/src/globals.cpp:39-116
    const char sep = ',';
    long double total = 0;
    const uint_t n = (uint_t)flags.n;
    for (uint_t i = 0; i < n + h + 1; ++i) total += static_cast<long double>(counts[i]);
    long double base1 = pow(2.0, n);
    const int cnts = (int)flags.n + (int)h + 1;
    if (*id == 0) {
        ss << "id, k, A, h, A_2, m_1, m_2, hlen";
        for (int i = 1; i <= cnts; ++i) ss << ", p_" << i;
        for (int i = 1; i <= cnts; ++i) ss << ", hist_" << i;
        ss << ", t_kernel, t, sum, minb, mincb, superA, minb2, mincb2, superA2\n";
    }
    ss << (*id)++ << sep << flags.n << sep << A << sep << h << sep << printbits_str(A, n) << sep
       << flags.mc_iterations << sep << flags.mc_iterations_2 << sep << cnts;
    for (uint_t i = 1; i <= (uint_t)cnts; ++i) {
        const long double base2 = binom(static_cast<long double>(n + h), static_cast<long double>(i));
        const long double base = base1 * base2;
        const long double prob = counts[i] / base;
        ss << sep << prob;
    }
    for (uint_t i = 1; i <= (uint_t)cnts; ++i) {
        if (n >= 32 && counts[i] > (uint128_t(1) << 64))
            ss << sep << setprecision(12) << static_cast<long double>(counts[i]);
        else
            ss << sep << static_cast<uintll_t>(counts[i]);
    }
    for (int i = 0; i < 2; ++i) ss << sep << times[i];
    ss << sep << total;
    if (flags.verbose == 1) {
        std::cout << "Kernel: " << times[0] << " s\n"
                  << "Runtime: " << times[1] << " s\n"
                  << flags.n << sep << A << sep << h << sep << flags.mc_iterations << sep << flags.mc_iterations_2
                  << "\n";
    }
}

}  // namespace cr

void process_result(std::stringstream &ss, uintll_t A, uint_t h, const Flags &flags, double *times,
                    uint128_t *counts) {
    static unsigned int id = 0;
    cr::supera_row(ss, A, h, flags, times, counts, &id);
}
