/* Exhaustive check of the box smoothing's division (div_box in
 * kirch_datuming_amd/csrc/kmig_prep.hip): for d = 2 r + 1 and every finite
 * float s, the product with the reciprocal and one fused correction,
 *     q0 = s * (1/d),  e = fma(-q0, d, s),  q = fma(e, 1/d, q0),
 * equals the IEEE quotient s / d bit for bit, except for s = -0.0 (a box sum
 * is never -0.0).  Round-to-nearest-even is symmetric in sign, so both
 * sides give q(-s) = -q(s), and the non-negative floats cover all of them.
 * Prints "d=<d> mismatches <n> first <hex>" per divisor.
 * usage: div_exact d...   (built with gcc -O2 -march=native -fopenmp -ffp-contract=off) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv)
{
    for (int k = 1; k < argc; k++) {
        const float d = (float)atoi(argv[k]), r = 1.0f / d;
        long bad = 0;
        int64_t first = (int64_t)1 << 32;
#pragma omp parallel for reduction(+ : bad) reduction(min : first) schedule(static)
        for (int64_t i = 0; i < ((int64_t)1 << 31); i++) {
            const uint32_t u = (uint32_t)i;
            const float s = bits_f(u);
            if (!isfinite(s)) continue;
            const float q0 = s * r;
            const float e = fmaf(-q0, d, s);
            const float q = fmaf(e, r, q0);
            if (f_bits(q) != f_bits(s / d)) {
                bad++;
                if (i < first) first = i;
            }
        }
        printf("d=%s mismatches %ld first %08llx\n", argv[k], bad,
               (unsigned long long)(bad ? first : 0));
    }
    return 0;
}
