/* Exhaustive check of the fused tuple pass's division by a fixed divisor D
 * (div_fixed in kirch_datuming_amd/csrc/kmig_prep.hip): for every float x
 * with 2^-60 <= |x| (finite or infinite) and the divisor's double reciprocal
 * r = 1.0 / (double)D,
 *     (float)((double)x * r)  ==  x / D   (IEEE float division),
 * bit for bit.  (The double product is within 2^-52 of x/D relative; a
 * float quotient that is not a float midpoint lies at least 2^-49 from one,
 * and none is a midpoint, so both round alike.)  x below 2^-60 takes the
 * IEEE division in the kernel.  Rounding is symmetric in sign, so the
 * non-negative x cover all.  Prints "d=<D> mismatches <n>" per divisor; a
 * divisor written "f:<D>" uses the float reciprocal instead (a planted
 * mistake that must show mismatches).
 * usage: recip_exact D...   (gcc -O2 -fopenmp -ffp-contract=off) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv)
{
    const uint32_t lo = f_bits(0x1p-60f);
    for (int k = 1; k < argc; k++) {
        const int single = strncmp(argv[k], "f:", 2) == 0;
        const float d = strtof(argv[k] + (single ? 2 : 0), NULL);
        const double r = 1.0 / (double)d;
        const float rf = 1.0f / d;
        long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
        for (int64_t i = lo; i <= 0x7f800000LL; i++) {
            const float x = bits_f((uint32_t)i);
            const float q = single ? x * rf : (float)((double)x * r);
            if (f_bits(q) != f_bits(x / d)) bad++;
        }
        printf("d=%s mismatches %ld\n", argv[k], bad);
    }
    return 0;
}
