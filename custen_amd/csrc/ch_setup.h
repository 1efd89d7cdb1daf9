// ch_setup.h -- factorisation of the Cahn-Hilliard ADI matrix, shared by the
// host setup of every simulation and by the device-side setup of the
// allocation-free entry point, so that both produce the same bits.
//
// The matrix of every line solve is the cyclic pentadiagonal
// [a, b, c, d, e] = [s, -4s, 1+6s, -4s, s].  Its leading (n-2) block E is
// factored in place (serialCahnADI.c:91-155, cuPentBatch.cu:35-113); the
// corner columns give inv1 = E^-1 f1 and inv2 = E^-1 f2 with
// f1 = [a, 0, ..., e, d] and f2 = [b, a, 0, ..., e] (serialCahnADI.c:247-417,
// BatchHyper.cu:115-189), and omega is the inverse of the 2x2 Schur complement
// (BatchHyper.cu:265-514).  Every operation is rounded as written: the files
// that include this are compiled with -ffp-contract=off.
#pragma once

#if defined(__HIPCC__)
#define CUSTEN_CH_HD __host__ __device__
#else
#define CUSTEN_CH_HD
#endif

namespace custen {

CUSTEN_CH_HD inline void chPentFactor(double* __restrict__ ds, double* __restrict__ dl,
                                      double* __restrict__ d, double* __restrict__ du,
                                      double* __restrict__ dw, int m)
{
    // serialCahnADI.c:91-155 / cuPentBatch.cu:35-113
    int r = 0;
    du[r] = du[r] / d[r];
    dw[r] = dw[r] / d[r];
    int rp = r;
    r += 1;
    d[r] = d[r] - dl[r] * du[rp];
    du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    dw[r] = dw[r] / d[r];
    for (int i = 2; i < m - 2; i++) {
        const int rpp = r - 1;
        rp = r;
        r += 1;
        dl[r] = dl[r] - ds[r] * du[rpp];
        d[r] = d[r] - ds[r] * dw[rpp] - dl[r] * du[rp];
        dw[r] = dw[r] / d[r];
        du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    }
    for (int last = 0; last < 2; last++) {
        const int rpp = r - 1;
        rp = r;
        r += 1;
        dl[r] = dl[r] - ds[r] * du[rpp];
        d[r] = d[r] - ds[r] * dw[rpp] - dl[r] * du[rp];
        if (last == 0) du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    }
}

CUSTEN_CH_HD inline void chPentSolve(const double* __restrict__ ds,
                                     const double* __restrict__ dl, const double* __restrict__ d,
                                     const double* __restrict__ du, const double* __restrict__ dw,
                                     double* __restrict__ b, int m)
{
    // serialCahnADI.c:161-241 / cuPentBatch.cu:119-198
    b[0] = b[0] / d[0];
    b[1] = (b[1] - dl[1] * b[0]) / d[1];
    for (int i = 2; i < m; i++) b[i] = (b[i] - ds[i] * b[i - 2] - dl[i] * b[i - 1]) / d[i];
    b[m - 2] = b[m - 2] - du[m - 2] * b[m - 1];
    for (int i = m - 3; i >= 0; i--) b[i] = b[i] - du[i] * b[i + 1] - dw[i] * b[i + 2];
}

// The factored rows of a constant-coefficient matrix converge: from some row
// on they repeat bitwise.  fwd = the first row from which ds, dl, d equal
// row m-1's (the rows the forward substitution reads up to its last row);
// bwd = the first row from which du, dw equal row m-3's (the last row the
// backward substitution's steps read).  When the factors never settle (a
// two-value cycle, say) the runs are the last row alone.
CUSTEN_CH_HD inline void chConstantRuns(const double* fac, int m, int& fwd, int& bwd)
{
    const double* ds = fac;
    const double* dl = ds + m;
    const double* dd = dl + m;
    const double* du = dd + m;
    const double* dw = du + m;
    // bitwise: +0.0 and -0.0 are told apart, a NaN matches its own bits
    auto bits = [](double v) { return __builtin_bit_cast(unsigned long long, v); };
    const int lf = m - 1, lb = m - 3;
    fwd = lf;
    while (fwd > 0 && bits(ds[fwd - 1]) == bits(ds[lf]) && bits(dl[fwd - 1]) == bits(dl[lf]) &&
           bits(dd[fwd - 1]) == bits(dd[lf]))
        fwd--;
    bwd = lb;
    while (bwd > 0 && bits(du[bwd - 1]) == bits(du[lb]) && bits(dw[bwd - 1]) == bits(dw[lb])) bwd--;
}

// fac: 7 * m doubles, laid out ds, dl, d, du, dw, inv1, inv2; every entry is
// written here (m = n - 2 >= 6).
CUSTEN_CH_HD inline void chFactor(double* fac, int m, double a, double b, double c, double d,
                                  double e, double omega[4])
{
    double* ds = fac;
    double* dl = ds + m;
    double* dd = dl + m;
    double* du = dd + m;
    double* dw = du + m;
    double* inv1 = dw + m;
    double* inv2 = inv1 + m;
    for (int i = 0; i < m; i++) {
        ds[i] = i > 1 ? a : 0.0;
        dl[i] = i > 0 ? b : 0.0;
        dd[i] = c;
        du[i] = i < m - 1 ? d : 0.0;
        dw[i] = i < m - 2 ? e : 0.0;
        inv1[i] = 0.0;
        inv2[i] = 0.0;
    }
    chPentFactor(ds, dl, dd, du, dw, m);
    inv1[0] = a;
    inv1[m - 2] = e;
    inv1[m - 1] = d;
    inv2[0] = b;
    inv2[1] = a;
    inv2[m - 1] = e;
    chPentSolve(ds, dl, dd, du, dw, inv1, m);
    chPentSolve(ds, dl, dd, du, dw, inv2, m);
    const double z11 = e * inv1[0] + a * inv1[m - 2] + b * inv1[m - 1];
    const double z12 = e * inv2[0] + a * inv2[m - 2] + b * inv2[m - 1];
    const double z21 = d * inv1[0] + e * inv1[1] + a * inv1[m - 1];
    const double z22 = d * inv2[0] + e * inv2[1] + a * inv2[m - 1];
    const double y11 = c - z11, y12 = d - z12, y21 = b - z21, y22 = c - z22;
    const double det = 1.0 / (y11 * y22 - y21 * y12);
    omega[0] = y22 * det;
    omega[1] = -y12 * det;
    omega[2] = -y21 * det;
    omega[3] = y11 * det;
}

}  // namespace custen
