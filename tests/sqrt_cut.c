/* Exhaustive check of the root-free gradient test (grad_vanishes in
 * kirch_datuming_amd/csrc/kmig_prep.hip): for every non-negative float s
 * (finite, infinite or NaN; a sum of two squares is never negative),
 *     sqrtf(s) <= 1e-12f   ==   s <= S,
 * with S given as a bit pattern in hex.  Prints "cut=<hex> mismatches <n>".
 * usage: sqrt_cut hex...   (built with gcc -O2 -fopenmp -ffp-contract=off) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

int main(int argc, char **argv)
{
    for (int k = 1; k < argc; k++) {
        const float cut = bits_f((uint32_t)strtoul(argv[k], NULL, 16));
        long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
        for (int64_t i = 0; i < ((int64_t)1 << 31); i++) {
            const float s = bits_f((uint32_t)i);
            if ((sqrtf(s) <= 1.0e-12f) != (s <= cut)) bad++;
        }
        printf("cut=%s mismatches %ld\n", argv[k], bad);
    }
    return 0;
}
