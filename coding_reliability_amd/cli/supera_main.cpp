// Super-A search command line (distance_distribution_super_a/src/main.cpp:28-387)
// on the MI355X engine: every exhaustive histogram of the sweep in one
// batched launch.  Flags: -k bits, -s start, -S end, -c candidates file,
// -m/-M grid sizes, -f prefix, -v level, -i device.  Writes
// <prefix>_k<k>_a<start>-<end>.csv (or _c_<file>.csv) like the reference.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "coding_reliability.h"

int main(int argc, char **argv) {
    unsigned long long k = 16, start = 0, end = 0, m1 = 0, m2 = 0;
    int grid = 0, verbose = 0, dev = 0, mode = 0;
    const char *file = nullptr, *prefix = "result";
    for (int i = 0; i < argc; ++i) {
        auto nxt = [&]() { return i + 1 < argc ? argv[i + 1] : "0"; };
        if (!strcmp(argv[i], "-h")) {
            std::cout << "\nUSAGE:\n\n\t-k NUM\t number of bits as input size\n\t-s NUM\t search start A\n"
                         "\t-S NUM\t search end A\n\t-c FILE\t candidates file\n\t-m NUM\t 1-D grid size\n"
                         "\t-M NUM\t 2-D grid size\n\t-f STR\t output prefix\n\t-i NUM\t GPU index\n"
                         "\t-v NUM\t verbosity\n\n";
            return 0;
        }
        if (!strcmp(argv[i], "-v")) verbose = atoi(nxt());
        else if (!strcmp(argv[i], "-s")) { mode = 1; start = strtoull(nxt(), nullptr, 10); }
        else if (!strcmp(argv[i], "-S")) { mode = 1; end = strtoull(nxt(), nullptr, 10); }
        else if (!strcmp(argv[i], "-c")) { mode = 2; file = nxt(); }
        else if (!strcmp(argv[i], "-m")) { grid = 1; m1 = strtoull(nxt(), nullptr, 10); }
        else if (!strcmp(argv[i], "-M")) { grid = 2; m2 = strtoull(nxt(), nullptr, 10); }
        if (!strcmp(argv[i], "-f")) prefix = nxt();
        if (!strcmp(argv[i], "-k")) k = strtoull(nxt(), nullptr, 10);
        if (!strcmp(argv[i], "-i")) dev = atoi(nxt());
    }
    if (grid == 2 && m2 == 0) m2 = m1;
    if (start > end) end = start;
    if (start == 0 && end > 0) start = end;
    if (mode == 0) {
        std::cerr << "give a search range (-s/-S) or a candidates file (-c)\n";
        return 1;
    }
    std::vector<uint64_t> As;
    if (mode == 2) {
        // Comma-separated values per line (super_a/src/main.cpp:172-192).  The
        // ANcandidates output is accepted as well: '#' lines are comments and a
        // tab-separated row contributes its last column, the candidate list
        // (ANcandidates.cpp:286-299).
        std::ifstream fp(file);
        if (!fp.good()) {
            std::cerr << "Candidates file does not exist.\n";
            return 1;
        }
        std::string line, tok;
        while (std::getline(fp, line)) {
            if (!line.empty() && line[0] == '#') continue;
            const size_t tab = line.rfind('\t');
            if (tab != std::string::npos) line = line.substr(tab + 1);
            std::stringstream ss(line);
            while (std::getline(ss, tok, ',')) {
                const size_t b = tok.find_first_not_of(" \r");
                if (b == std::string::npos) continue;
                As.push_back(std::stoull(tok.substr(b)));
            }
        }
    }
    cr_context *ctx = nullptr;
    if (cr_context_create(dev, &ctx) != CR_OK) {
        std::cerr << cr_last_error() << std::endl;
        return 5;
    }
    cr_supera_request q{};
    q.style = 1;
    q.k = (uint32_t)k;
    q.with_grid = grid;
    q.A_start = start;
    q.A_end = end;
    q.As = As.empty() ? nullptr : As.data();
    q.nAs = (uint32_t)As.size();
    q.verbose = verbose;
    q.m1 = m1;
    q.m2 = m2;
    char *text = nullptr;
    if (cr_supera_search(ctx, &q, &text, nullptr) != CR_OK) {
        std::cerr << cr_last_error() << std::endl;
        cr_context_destroy(ctx);
        return 6;
    }
    char fname[512];
    if (mode == 1) {
        snprintf(fname, sizeof(fname), "%s_k%llu_a%llu-%llu.csv", prefix, k, start, end);
    } else {
        std::string tmp(file);
        std::replace(tmp.begin(), tmp.end(), '/', '_');
        snprintf(fname, sizeof(fname), "%s_k%llu_c_%s.csv", prefix, k, tmp.c_str());
    }
    std::ofstream(fname) << text;
    cr_free(text);
    cr_context_destroy(ctx);
    return 0;
}
