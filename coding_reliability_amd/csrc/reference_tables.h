// Golden tables of the reference (see reference_tables.cpp).
#pragma once

namespace cr {
namespace tables {
extern const unsigned long long sol_hamming_n8[];
extern const unsigned sol_hamming_n8_len;
extern const unsigned long long sol_hamming_n16[];
extern const unsigned sol_hamming_n16_len;
extern const unsigned long long sol_hamming_n24[];
extern const unsigned sol_hamming_n24_len;
extern const unsigned long long sol_hamming_n32[];
extern const unsigned sol_hamming_n32_len;
extern const unsigned long long solution_an8_A61[];
extern const unsigned solution_an8_A61_len;
extern const unsigned long long solution_an16_A61[];
extern const unsigned solution_an16_A61_len;
extern const unsigned long long solution_an24_A61[];
extern const unsigned solution_an24_A61_len;
extern const unsigned long long solution_an24_A641[];
extern const unsigned solution_an24_A641_len;
extern const unsigned long long solution_an16_A641[];
extern const unsigned solution_an16_A641_len;
extern const unsigned long long solution_an8_A641[];
extern const unsigned solution_an8_A641_len;
}  // namespace tables
}  // namespace cr
