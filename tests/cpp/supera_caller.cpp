// Test caller of the super_a C++ entry points in libcodrel.so, compiled
// against include/cr_reference_api.h as a reference driver would be:
//   run_ancoding(A, h, Flags&, times, counts)       distance_distribution_super_a/src/algorithms.h:20
//   run_ancoding_grid(A, h, Flags&, times, counts)  distance_distribution_super_a/src/algorithms.h:21
//   process_result(ss, A, h, flags, times, counts)  distance_distribution_super_a/src/globals.h:58-62
// It drives them the way super_a's main.cpp does (h = floor(log2 A) + 1,
// main.cpp:209,261) and prints one line per call for a test to check:
//   R <A> <h> <t_kernel> <t_total> <counts[0]> ... <counts[k+h]>
// followed by the process_result CSV text after a line "CSV".
//
// usage: supera_caller exhaustive <k> <dev> <A>...
//        supera_caller grid <gdim> <k> <m1> <m2> <A>...
// dev < 0 leaves Flags::dev at its default (-1).  A call that throws prints
// "E <A> <message>" and the program goes on with the next A.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "cr_reference_api.h"

static std::string u128_str(uint128_t v) {
    std::ostringstream os;
    os << v;  // the library's operator<< for uint128_t (globals.cpp:320-341)
    return os.str();
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s exhaustive <k> <dev> <A>... | grid <gdim> <k> <m1> <m2> <A>...\n", argv[0]);
        return 2;
    }
    const bool grid = strcmp(argv[1], "grid") == 0;
    Flags flags;
    int first;
    if (grid) {
        if (argc < 7) return 2;
        flags.with_grid = atoi(argv[2]);
        flags.n = strtoull(argv[3], nullptr, 10);
        flags.mc_iterations = atoi(argv[4]);
        flags.mc_iterations_2 = atoi(argv[5]);
        first = 6;
    } else {
        flags.n = strtoull(argv[2], nullptr, 10);
        const int dev = atoi(argv[3]);
        if (dev >= 0) flags.dev = dev;
        first = 4;
    }
    std::stringstream ss;
    for (int a = first; a < argc; ++a) {
        const uintll_t A = strtoull(argv[a], nullptr, 10);
        const uint_t h = (uint_t)(floor(log((double)A) / log(2.0)) + 1);
        uint128_t counts[COUNTS_MAX_WIDTH] = {0};
        double times[2] = {0, 0};
        try {
            if (grid)
                run_ancoding_grid(A, h, flags, times, counts);
            else
                run_ancoding(A, h, flags, times, counts);
        } catch (const std::exception &e) {
            std::cout << "E " << A << " " << e.what() << "\n";
            continue;
        }
        std::cout << "R " << A << " " << h << " " << times[0] << " " << times[1];
        for (uint_t i = 0; i < flags.n + h + 1 && i < COUNTS_MAX_WIDTH; ++i) std::cout << " " << u128_str(counts[i]);
        std::cout << "\n";
        process_result(ss, A, h, flags, times, counts);
        ss << "\n";
    }
    std::cout << "CSV\n" << ss.str();
    return 0;
}
