/*
 * Thin C entry points over the reference's own sqaod::Random
 * (sqaodc/common/Random.cpp, compiled from the reference tree by
 * oracle/Makefile into oracle/_ref).  Test infrastructure only: used to pin
 * the oracle's MT19937 restatement against the real code.
 */
#include "Random.h"
#include <stddef.h>
#include <stdint.h>

extern "C" {

void sqref_mt_stream(unsigned long long seed, unsigned long long skip, uint32_t *out, size_t n) {
    sqaod::Random r;
    r.seed(seed);
    for (unsigned long long i = 0; i < skip; ++i)
        (void)r.randInt32();
    for (size_t i = 0; i < n; ++i)
        out[i] = (uint32_t)r.randInt32();
}

void sqref_mt_f32(unsigned long long seed, float *out, size_t n) {
    sqaod::Random r;
    r.seed(seed);
    for (size_t i = 0; i < n; ++i)
        out[i] = r.random<float>();
}

void sqref_mt_f64(unsigned long long seed, double *out, size_t n) {
    sqaod::Random r;
    r.seed(seed);
    for (size_t i = 0; i < n; ++i)
        out[i] = r.random<double>();
}

void sqref_mt_randint(unsigned long long seed, int high, uint32_t *out, size_t n) {
    sqaod::Random r;
    r.seed(seed);
    for (size_t i = 0; i < n; ++i)
        out[i] = (uint32_t)r.randInt(high);
}

}
