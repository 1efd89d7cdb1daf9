// fastdiv_check.cpp -- the divisor check of the two-operation division
// (chdiv::certified, custen_amd/csrc/ch_setup.h) against brute force.
//
// float: for each divisor every significand of x in [1, 2) is tried (2^23
// values), so "the division is correctly rounded for every x" is known
// exactly; the check's answer must agree with it in both directions (a
// certified divisor with a failing x would make the fast path wrong; a
// refused divisor without one would mean the list of candidates is not the
// failing set).
// double: the divisors of the Cahn-Hilliard column solve at several grid
// sizes and random divisors; each certified divisor is tried on random x of
// every exponent the kernels allow and on the x next to its candidates.
//
// usage: fastdiv_check <float divisors> <double divisors> <x per double divisor> <seed>
//        [narrow]   (narrow: the float check lists no candidates)
// prints "float <n> certified <c> disagree <k>" and
// "double <n> certified <c> mismatches <k>"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../custen_amd/csrc/ch_setup.h"

namespace {

uint64_t state;
uint64_t next()  // SplitMix64
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float fbits(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

double dbits(uint64_t u)
{
    double f;
    std::memcpy(&f, &u, 8);
    return f;
}

// every x in [1, 2): is fma(x, zh, RN(x zl)) == x / d?
bool floatExhaustive(float d)
{
    const float zh = 1.0f / d;
    const float zl = std::fmaf(-d, zh, 1.0f) / d;
    for (uint32_t m = 0; m < (1u << 23); m++) {
        const float x = fbits(0x3f800000u | m);
        if (!(std::fmaf(x, zh, x * zl) == x / d)) return false;
    }
    return true;
}

// the CH matrix's constant-run pivot for grid n (the bench's coefficients)
double chPivot(int n, double gamma)
{
    const double lx = 2 * M_PI, dx = lx / n, dt = 0.1 * dx;
    const double s = 2.0 * dt * 1.0 * gamma / (3.0 * (dx * dx * dx * dx));
    const int m = n - 2;
    double* fac = new double[(size_t)7 * m];
    double om[4];
    custen::chFactor(fac, m, s, -4 * s, 1 + 6 * s, -4 * s, s, om);
    int f = 0, b = 0;
    custen::chConstantRuns(fac, m, f, b);
    const double piv = fac[2 * m + (f < m - 1 ? f : m - 1)];
    delete[] fac;
    return piv;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    const int nf = std::atoi(argv[1]), nd = std::atoi(argv[2]), nx = std::atoi(argv[3]);
    state = std::strtoull(argv[4], nullptr, 10);
    const bool narrow = argc > 5 && std::strcmp(argv[5], "narrow") == 0;

    // float, both directions
    int fcert = 0, disagree = 0;
    const uint32_t special[] = {0, 1, 2, 3, 0x400000, 0x200000, 0x600000, 0x7fffff, 0x555555,
                                0x2aaaaa, 0x124925, 0x49249};
    const int ns = (int)(sizeof special / sizeof special[0]);
    for (int i = 0; i < nf; i++) {
        const uint32_t m = i < ns ? special[i] : (uint32_t)(next() & 0x7fffff);
        const float d = fbits(0x3f800000u | m);
        const bool c = narrow ? custen::chdiv::certified<float, uint32_t, 24, 0>(d)
                              : custen::chdiv::certified<float, uint32_t, 24>(d);
        const bool e = floatExhaustive(d);
        fcert += c;
        if (c != e) {
            disagree++;
            std::printf("disagree d=%a certified=%d exhaustive=%d\n", (double)d, c, e);
        }
    }
    std::printf("float %d certified %d disagree %d\n", nf, fcert, disagree);

    // double: CH pivots and random divisors, sampled x
    int dcert = 0, tried = 0;
    long mism = 0;
    const int sizes[] = {64, 128, 130, 256, 512, 1024, 2048, 4096};
    const int nsz = (int)(sizeof sizes / sizeof sizes[0]);
    for (int i = 0; i < nsz + nd; i++) {
        double d;
        if (i < nsz) {
            d = chPivot(sizes[i], 0.01);
            std::printf("pivot n %d d %.17g certified %d\n", sizes[i], d,
                        (int)custen::chDivCertified(d));
        } else {
            const int ex = (int)(next() % 121) - 60;
            d = std::ldexp(dbits(0x3ff0000000000000ull | (next() & 0xfffffffffffffull)), ex);
            if (next() & 1) d = -d;
        }
        tried++;
        if (!custen::chDivCertified(d)) continue;
        dcert++;
        const double zh = 1.0 / d;
        const double zl = std::fma(-d, zh, 1.0) / d;
        for (int k = 0; k < nx; k++) {
            const int ex = (int)(next() % 1801) - 900;
            double x = std::ldexp(dbits(0x3ff0000000000000ull | (next() & 0xfffffffffffffull)), ex);
            if (next() & 1) x = -x;
            if (!(std::fma(x, zh, x * zl) == x / d)) mism++;
        }
    }
    std::printf("double %d certified %d mismatches %ld\n", tried, dcert, mism);
    return 0;
}
