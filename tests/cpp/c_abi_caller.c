/* A plain C (C99) caller of the C ABI in include/coding_reliability.h: one
 * exhaustive AN histogram, one Extended Hamming weight histogram, two AN
 * Monte Carlo sample ranges and one refused call, printed one result per line
 * for tests/test_gpu_c_abi.py to compare with the oracle. */
#include <stdint.h>
#include <stdio.h>

#include "coding_reliability.h"

static void print_row(const char *name, const uint64_t *v, unsigned n) {
    printf("%s", name);
    for (unsigned i = 0; i < n; ++i) printf(" %llu", (unsigned long long)v[i]);
    printf("\n");
}

int main(void) {
    cr_context *ctx = NULL;
    cr_status s = cr_context_create(0, &ctx);
    if (s != CR_OK) {
        fprintf(stderr, "cr_context_create: %s (%s)\n", cr_status_string(s), cr_last_error());
        return 2;
    }
    /* exhaustive AN, A = 61, 2^8 words: raw unordered-pair counts, 15 bins */
    uint64_t raw[15];
    double ms = 0;
    s = cr_an_pairs(ctx, 61, 256, raw, 15, &ms);
    if (s != CR_OK) {
        fprintf(stderr, "cr_an_pairs: %s\n", cr_last_error());
        return 3;
    }
    print_row("an_pairs", raw, 15);
    /* Extended Hamming, k = 16: weights of all 2^16 non-extended codewords */
    cr_ham_request req = {16, CR_HAM_RANGE, 0, 65536, 0, NULL, 0, 0, 0};
    uint64_t w[23];
    s = cr_hamming_weights(ctx, &req, w, 23, &ms);
    if (s != CR_OK) {
        fprintf(stderr, "cr_hamming_weights: %s\n", cr_last_error());
        return 4;
    }
    print_row("hamming16", w, 23);

    /* the same request and a (13,8) window in one batched call */
    cr_ham_request batch[2] = {{16, CR_HAM_RANGE, 0, 65536, 0, NULL, 0, 0, 0},
                               {8, CR_HAM_RANGE, 100, 100, 0, NULL, 0, 0, 0}};
    uint64_t bw[2 * 23];
    s = cr_hamming_weights_batch(ctx, batch, 2, bw, 23, &ms);
    if (s != CR_OK) {
        fprintf(stderr, "cr_hamming_weights_batch: %s\n", cr_last_error());
        return 6;
    }
    print_row("batch_hamming16", bw, 23);
    print_row("batch_hamming8_100_200", bw + 23, 23);
    /* AN Monte Carlo, A = 61, n = 10, seed 7: sample ranges [0, 100) and [100, 300) */
    const uint64_t bounds[3] = {0, 100, 300};
    uint64_t hists[2 * 65];
    s = cr_an_mc_ranges(ctx, 61, 10, bounds, 2, 7, hists, 65, &ms);
    if (s != CR_OK) {
        fprintf(stderr, "cr_an_mc_ranges: %s\n", cr_last_error());
        return 5;
    }
    print_row("an_mc_0_100", hists, 65);
    print_row("an_mc_100_300", hists + 65, 65);
    /* a refused call: A = 0 */
    s = cr_an_pairs(ctx, 0, 256, raw, 15, NULL);
    printf("refused %d %s\n", (int)s, cr_last_error());
    cr_context_destroy(ctx);
    return 0;
}
