/* division_check.c -- the reciprocal-based division of the fused CH column
 * solve, restated on the CPU: q = RN(n y) with y = RN(1/d), then two
 * corrections q <- RN(q + RN(n - d q) y) (remainders by fma).  For random
 * operands with 2^-900 <= |n| <= 2^900 and 2^-60 <= |d| <= 2^60 it counts the
 * quotients that differ from the IEEE quotient n / d in any bit.
 *
 * usage: division_check <samples> <seed>   prints "<samples> <mismatches>" */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t state;
static uint64_t next(void) /* SplitMix64 */
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* a double with a random sign and 52-bit mantissa and exponent in [lo, hi] */
static double draw(int lo, int hi)
{
    const uint64_t r = next();
    const int e = lo + (int)(next() % (uint64_t)(hi - lo + 1));
    const double m = 1.0 + (double)(r >> 12) * 0x1p-52;
    const double v = ldexp(m, e);
    return (r & 1) ? -v : v;
}

static double divByRecip(double n, double d, double y)
{
    double q = n * y;
    q = fma(fma(-d, q, n), y, q);
    return fma(fma(-d, q, n), y, q);
}

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    const long samples = atol(argv[1]);
    state = strtoull(argv[2], NULL, 10);
    long bad = 0;
    for (long i = 0; i < samples; i++) {
        /* the range ends and the middle alike */
        const int k = (int)(i % 4);
        const double n = k == 0 ? draw(-900, -890) : (k == 1 ? draw(890, 899) : draw(-60, 60));
        const double d = k == 2 ? draw(-60, -55) : (k == 3 ? draw(55, 59) : draw(-60, 59));
        const double y = 1.0 / d;
        const double q = divByRecip(n, d, y);
        const double ref = n / d;
        if (memcmp(&q, &ref, sizeof q) != 0) bad++;
    }
    printf("%ld %ld\n", samples, bad);
    return 0;
}
