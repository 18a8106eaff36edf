/*
 * cr_reference_api.h -- the reference's C++ entry points, implemented on the
 * MI355X engine (libcodrel.so).  Signatures, argument meaning and output
 * formats are those of brics-db/coding_reliability, so its drivers relink
 * unchanged:
 *   distance_distribution/src/algorithms.h:20-34   run_* entry points
 *   distance_distribution/src/globals.h:34-47      error and result-printing helpers
 *   distance_distribution_super_a/src/globals.h:28-44, :58-62, algorithms.h:20-21
 *                                                  Flags, run_ancoding / run_ancoding_grid, process_result
 * Differences from the reference are limited to what the hardware changes:
 * one GPU (nr_dev and Flags::nr_dev/dev select device 0 / the given index),
 * no cudaDeviceReset, HIP-event kernel timing, Monte Carlo words from
 * Philox4x32-10 instead of cuRAND, errors reported as std::runtime_error
 * instead of exit().  The entry points of the reference's CPU implementations
 * (run_ancoding_cpu, run_hamming_cpu, run_hamming_cpu_native*) compute the
 * same numbers on the GPU and keep those paths' own conventions (no
 * power-of-two mapping of A, their bin counts, file prefixes and one-row
 * timing tables).  The cuRAND dumps (test_curand*) are not part of this
 * library: calling them throws std::runtime_error.
 */
#ifndef CR_REFERENCE_API_H_
#define CR_REFERENCE_API_H_

#include <cstddef>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

typedef unsigned uint_t;
typedef unsigned long long uintll_t;
typedef unsigned __int128 uint128_t;

/* Timing table printed by the result writers: (label, unit, value) rows whose
 * values are scaled by a common factor (the drivers record ms and print s).
 * Mirrors the parts of the helper library's Statistics the writers use. */
class Statistics {
  public:
    int add(const std::string &label, const std::string &unit);
    void set(int i, double value);
    void setFactorAll(double f) { factor_ = f; }
    int getLength() const { return (int)rows_.size(); }
    std::string getLabel(int i) const { return rows_.at(i).label; }
    std::string getUnit(int i) const { return rows_.at(i).unit; }
    double getAverage(int i) const {
        check_index(i);
        return rows_[i].value * factor_;
    }

  private:
    // As the helper library's Statistics::check_index
    // (distance_distribution/lib/helper/src/statistics.cpp:165-170): a message
    // on stderr, then std::out_of_range.
    void check_index(int i) const;
    struct Row {
        std::string label, unit;
        double value = 0;
    };
    std::vector<Row> rows_;
    double factor_ = 1.0;
};

/* distance_distribution/src/algorithms.h:20-34 */
void run_hamming(uintll_t n, int with_1bit, int file_output, int nr_dev);
void run_hamming_cpu(uintll_t n, int with_1bit, int file_output);
void run_hamming_cpu_native(uintll_t n, int with_1bit, int file_output);
void run_hamming_cpu_native_short(uintll_t n, int with_1bit, int file_output);
void run_hamming_cpu_native_short_ordering(uintll_t n, int with_1bit, int file_output);
void run_ancoding(uintll_t n, uintll_t A, int verbose, double *times, uintll_t *minb, uint128_t *mincb,
                  int file_output, int nr_dev);
void run_ancoding_cpu(uintll_t n, uintll_t A, int verbose, double *times, uintll_t *minb, uint128_t *mincb,
                      int file_output);
double run_hamming_mc(uintll_t n, int with_1bit, uintll_t iterations, int file_output, int nr_dev);
double run_ancoding_mc(uintll_t n, uintll_t iterations, uintll_t A, int verbose, double *times, uintll_t *minb,
                       uint128_t *mincb, int file_output, int nr_dev);
double run_ancoding_grid(int gdim, uintll_t n, uintll_t iterations, uintll_t iterations2, uintll_t A, int verbose,
                         double *times, uintll_t *minb, uint128_t *mincb, int file_output, int nr_dev);
double run_hamming_grid(uintll_t n, int with_1bit, uintll_t iterations, int file_output, int nr_dev);
void test_curand(uintll_t n, uintll_t iterations, int nr_dev);
void test_curand_1d(uintll_t n, uintll_t iterations, int nr_dev);
void test_curand_raw(uintll_t n, uintll_t iterations, int nr_dev);

/* distance_distribution/src/globals.h:34-47 */
double get_rel_error_hamming(uint_t n, uint128_t *tgt, int offset, int w1bit, double *errors = nullptr);
double get_abs_error_hamming(uint_t n, uint128_t *tgt, int offset, int w1bit, double *errors = nullptr);
double get_abs_error_AN(uintll_t A, uint_t n, uint128_t *tgt, int offset, double *errors = nullptr);
double get_rel_error_AN(uintll_t A, uint_t n, uint128_t *tgt, int offset, double *errors = nullptr);
void process_result_hamming(uint128_t *counts, Statistics stats, uint_t n, uint_t h, const char *file_prefix = nullptr);
void process_result_hamming_mc(uint128_t *counts, Statistics stats, uint_t n, uint_t h, int with_1bit,
                               uint_t iterations, const char *file_prefix = nullptr);
void process_result_ancoding(uint128_t *counts, Statistics stats, uint_t n, uint_t A, const char *file_prefix = nullptr,
                             std::string app = "");
void process_result_ancoding_mc(uint128_t *counts, Statistics stats, uint_t n, uint_t A, uint_t iterations,
                                const char *file_prefix = nullptr, std::string app = "");
std::ostream &operator<<(std::ostream &dest, uint128_t value);

/* distance_distribution_super_a/src/globals.h:28-44 */
struct Flags {
    uintll_t A = 61;
    int with_grid = 0;       // with "monte-carlo" grid approximation (=dim={1,2})
    int mc_iterations = 0;   // number of monte carlo iterations
    int mc_iterations_2 = 0; // for 2D-grid
    uintll_t n = 16;         // nr of bits as input size
    int search_super_A = 0;
    uintll_t search_start = 0;
    uintll_t search_end = 0;
    int nr_dev = 0;
    int dev = -1;
    int verbose = 0;
    const char *search_file = nullptr;
    const char *file_prefix = nullptr;
};
#ifndef COUNTS_MAX_WIDTH
#define COUNTS_MAX_WIDTH 64
#endif

/* distance_distribution_super_a/src/algorithms.h:20-21 */
void run_ancoding(uintll_t A, uint_t h, Flags &flags, double *times, uint128_t *counts);
void run_ancoding_grid(uintll_t A, uint_t h, Flags &flags, double *times, uint128_t *counts);

/* distance_distribution_super_a/src/globals.h:58-62 (one CSV row per code) */
void process_result(std::stringstream &ss, uintll_t A, uint_t h, const Flags &flags, double *times,
                    uint128_t *counts);

#endif /* CR_REFERENCE_API_H_ */
