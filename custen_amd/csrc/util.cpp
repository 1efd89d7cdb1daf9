// util.cpp -- host utilities exported by the library.
#include <cstddef>
#include <cstdint>

#include "../../include/cuSten.h"

#pragma STDC FP_CONTRACT OFF

extern "C" void custenFillUniform(double* out, size_t n, uint64_t seed, double a, double b)
{
    // SplitMix64 (Steele, Lea, Flood 2014): state_i = seed + (i+1) * golden gamma
    const double span = b - a;
    for (size_t i = 0; i < n; i++) {
        uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        const double u = (double)(z >> 11) * 0x1.0p-53;
        const double t = span * u;
        out[i] = a + t;
    }
}
