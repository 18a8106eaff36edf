// ANcandidates command line (ANcandidates/ANcandidates.cpp:171-301) on the
// MI355X engine: -a|--Amin -A|--Amax -d|--Dmin -D|--Dmax; one TSV line per
// (data width, A width) cell, candidates in increasing order.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "coding_reliability.h"

static void help(const char *argv0) {
    std::string b(argv0);
    size_t p = b.find_last_of('/');
    std::cout << "USAGE: " << b.substr(p == std::string::npos ? 0 : p)
              << " -a|--Amin -A|--Amax -d|--Dmin -D|--Dmax\n"
                 "\n\t-? -h --help               Print this message."
                 "\n\t-a --Amin                  Minimum width of A [bits]."
                 "\n\t-A --Amax                  Maximum width of A [bits]."
                 "\n\t-d --Dmin                  Minimum data width [bits]."
                 "\n\t-D --Dmax                  Maximum data width [bits]."
              << std::endl;
}

int main(int argc, char **argv) {
    size_t amin = 0, amax = 0, dmin = 0, dmax = 0;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        auto num = [&](size_t &dst, int code) -> int {
            if (i + 1 >= argc) {
                std::cerr << "argument '" << a << "' requires an additional number!" << std::endl;
                return code;
            }
            dst = std::strtoull(argv[++i], nullptr, 0);
            return 0;
        };
        int r = 0;
        if (!strcmp(a, "-h") || !strcmp(a, "--help") || !strcmp(a, "-?")) {
            help(argv[0]);
            return 255;
        } else if (!strcmp(a, "--Amin") || !strcmp(a, "-a")) {
            r = num(amin, 1);
        } else if (!strcmp(a, "--Amax") || !strcmp(a, "-A")) {
            r = num(amax, 2);
        } else if (!strcmp(a, "--Dmin") || !strcmp(a, "-d")) {
            r = num(dmin, 3);
        } else if (!strcmp(a, "--Dmax") || !strcmp(a, "-D")) {
            r = num(dmax, 4);
        }
        if (r) return r;
    }
    if (!amin || !amax || !dmin || !dmax) {
        help(argv[0]);
        return 100;
    }
    cr_context *ctx = nullptr;
    if (cr_context_create(0, &ctx) != CR_OK) {
        std::cerr << cr_last_error() << std::endl;
        return 5;
    }
    std::cout << "# bit width data = [" << dmin << ',' << dmax << "]\n";
    std::cout << "# bit width A = [" << amin << ',' << amax << "]\n";
    std::cout << "# n=|C|\tk=|D|\t|A|\t|SDR|\t|candidates|\tcandidates\n";
    const size_t nd = dmax - dmin + 1, na = amax - amin + 1;
    std::vector<uint64_t> cands(1u << 20);
    for (size_t cell = 0; cell < nd * na; ++cell) {
        const size_t k = cell % nd + dmin, ab = cell / nd + amin;
        uint64_t best = 0, n = 0;
        if (cr_an_candidates(ctx, (uint32_t)k, (uint32_t)ab, &best, cands.data(), cands.size(), &n, nullptr) != CR_OK) {
            std::cerr << cr_last_error() << std::endl;
            cr_context_destroy(ctx);
            return 6;
        }
        std::stringstream ss;
        ss << (k + ab) << '\t' << k << '\t' << ab << '\t' << best << '\t' << n << '\t';
        for (uint64_t i = 0; i < n && i < cands.size(); ++i) ss << (i ? "," : "") << cands[i];
        ss << '\n';
        std::cout << ss.str();
    }
    cr_context_destroy(ctx);
    return 0;
}
