// SDC tables of the Extended Hamming codes for a range of data widths, in the
// layout of the legacy table program (legacy/hamming_count_undetectable_errors.cpp:
// 226-339, one table per countHammingUndetectableErrors<k> call).
// Flags: -k first width, -K last width (default: first), -m method (0 auto,
// 1 device enumeration for k <= 40, 2 MacWilliams identity), -d device.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "coding_reliability.h"

int main(int argc, char **argv) {
    unsigned first = 0, last = 0;
    int method = 0, dev = 0;
    for (int i = 1; i < argc; ++i) {
        const char *v = i + 1 < argc ? argv[i + 1] : "0";
        if (!strcmp(argv[i], "-h")) {
            std::cout << "\nUSAGE:\n\n\t-k NUM\t first data width (1..64)\n\t-K NUM\t last data width\n"
                         "\t-m NUM\t 0 auto, 1 device enumeration (k <= 40), 2 MacWilliams identity\n"
                         "\t-d NUM\t GPU index\n\n";
            return 0;
        }
        if (!strcmp(argv[i], "-k")) first = (unsigned)strtoul(v, nullptr, 10);
        else if (!strcmp(argv[i], "-K")) last = (unsigned)strtoul(v, nullptr, 10);
        else if (!strcmp(argv[i], "-m")) method = atoi(v);
        else if (!strcmp(argv[i], "-d")) dev = atoi(v);
    }
    if (last < first) last = first;
    if (first < 1 || last > 64) {
        std::cerr << "data widths must lie in [1, 64]\n";
        return 1;
    }
    cr_context *ctx = nullptr;
    const bool device = method == 1 || (method == 0 && first <= 40);
    if (device && cr_context_create(dev, &ctx) != CR_OK) {
        std::cerr << cr_last_error() << std::endl;
        return 5;
    }
    int rc = 0;
    for (unsigned k = first; k <= last && rc == 0; ++k) {
        char *text = nullptr;
        if (cr_hamming_sdc_table(ctx, k, method, nullptr, 0, &text) != CR_OK) {
            std::cerr << cr_last_error() << std::endl;
            rc = 6;
            break;
        }
        std::cout << text << std::flush;
        cr_free(text);
    }
    if (ctx) cr_context_destroy(ctx);
    return rc;
}
