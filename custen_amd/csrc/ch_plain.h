// ch_plain.h -- reference-shaped Cahn-Hilliard step (the plain baseline).
#pragma once

#include <hip/hip_runtime_api.h>

#include <functional>

namespace custen {

struct PlainCH {
    int n = 0;
    double a = 0, b = 0, d = 0, e = 0;
    double omega[4] = {0, 0, 0, 0};
    double* cBar = nullptr;    // n*n
    double* nonLin = nullptr;  // n*n
    double* cHalf = nullptr;   // n*n
    double* multi = nullptr;   // 7 * (n-2) * n: ds, dl, d, du, dw, inv1, inv2 replicated per system
};

// One step in place on (cCurr, cOld).  `stencil(out, in, nonlinear)` applies
// the 3x3 u^3-u Laplacian (nonlinear) or the 5x5 biharmonic through the cuSten
// API.
hipError_t chPlainStep(const PlainCH& p, double* cCurr, double* cOld, hipStream_t st,
                       const std::function<hipError_t(double* out, const double* in, bool nonlinear)>& stencil);

}  // namespace custen
