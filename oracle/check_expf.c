/*
 * Exhaustive check: sq_expf (sqaod_amd/csrc/sq_expf.h, the device's exp)
 * against the host C library's expf, over all 2^32 float bit patterns.
 * This pins the device acceptance test to the CPU annealer's
 * `std::exp(-dE * beta)` (CPUDenseGraphAnnealer.cpp:262).
 *
 * Test infrastructure only (part of the oracle).  Build:
 *   gcc -O2 -ffp-contract=off -fopenmp check_expf.c -lm
 * Usage: check_expf [lo_bits hi_bits]   (default: whole range)
 * Exit status 0 iff every non-NaN result is bit-identical and NaN maps to NaN.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../sqaod_amd/csrc/sq_expf.h"

int main(int argc, char **argv) {
    uint64_t lo = 0, hi = 0x100000000ull;
    if (argc >= 3) {
        lo = strtoull(argv[1], NULL, 0);
        hi = strtoull(argv[2], NULL, 0);
    }
    long long mismatches = 0;
    uint32_t firstBad = 0;
    int haveBad = 0;
#pragma omp parallel for schedule(static, 1 << 20) reduction(+ : mismatches)
    for (long long b = (long long)lo; b < (long long)hi; ++b) {
        uint32_t bits = (uint32_t)b;
        float x;
        memcpy(&x, &bits, 4);
        float ref = expf(x);
        float got = sq_expf(x);
        uint32_t rb, gb;
        memcpy(&rb, &ref, 4);
        memcpy(&gb, &got, 4);
        int ok = (rb == gb) || (isnan(ref) && isnan(got));
        if (!ok) {
            ++mismatches;
#pragma omp critical
            {
                if (!haveBad) { haveBad = 1; firstBad = bits; }
            }
        }
    }
    printf("range [0x%llx, 0x%llx): %lld mismatches", (unsigned long long)lo,
           (unsigned long long)hi, mismatches);
    if (haveBad) {
        float x;
        memcpy(&x, &firstBad, 4);
        printf(" (first at bits 0x%08x x=%a: host %a, sq %a)", firstBad, x, expf(x), sq_expf(x));
    }
    printf("\n");
    return mismatches == 0 ? 0 : 1;
}

