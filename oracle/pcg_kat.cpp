// Prints PCG32 known answers from the reference's vendored PCG library (libs/pcg/pcg_random.hpp).
// Test infrastructure only: the output pins oracle.cpp's restatement of pcg32 (tests/golden/).
// Usage: pcg_kat <initstate> <initseq> <count>
#include <cstdio>
#include <cstdlib>
#include "pcg_random.hpp"

int main(int argc, char** argv) {
    unsigned long long st = argc > 1 ? std::strtoull(argv[1], nullptr, 0) : 42ull;
    unsigned long long sq = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 54ull;
    int n = argc > 3 ? std::atoi(argv[3]) : 16;
    pcg32 rng(st, sq);
    for (int i = 0; i < n; ++i) std::printf("%u\n", (unsigned)rng());
    return 0;
}
