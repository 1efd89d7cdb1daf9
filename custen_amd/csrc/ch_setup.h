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

// ---------------------------------------------------------------------------
// Division by a known divisor in two operations (Brisebarre, Muller, Raina,
// "Accelerating correctly rounded floating-point division when the divisor is
// known in advance", IEEE Trans. Computers 53(8), 2004): with zh = RN(1/d) and
// zl = RN((1 - d zh) / d) (the remainder 1 - d zh is exact by fma),
// q = RN(x zh + RN(x zl)) is RN(x / d) for every x unless x / d lies very
// close to a rounding boundary.  For a given d the x that come close enough
// can be listed, and chDivCertified tries each of them.
//
// Why the list is complete (precision p; |d| scaled to [1, 2), significand
// D = d 2^(p-1); x in [1, 2), significand X): |zh + zl - 1/d| <= 2^-(2p+2)
// and |RN(x zl) - x zl| <= 2^-(2p), so the sum before the last rounding is
// within E = 1.5 2^-(2p) of t = x / d.  q differs from RN(t) only when a
// midpoint mu between two neighbouring floats lies within E of t.  For t in
// [1, 2) those are mu = M 2^-p, for t in [1/2, 1) mu = M 2^-(p+1), with M odd
// in [2^p, 2^(p+1)); then |t - mu| = |2^P X - M D| / (2^P D) with P = p or
// p + 1, and the integer e = 2^P X - M D is never 0 (2^P X is even, M D is
// not a multiple of 2^P).  |e| <= E 2^P D < 3, so every failing x solves
// 2^P X - M D = e for some 0 < |e| <= 2; the check takes |e| <= 16.  With
// D = 2^s Do (Do odd), e must be 2^s times an odd number and
// M = -(e / 2^s) Do^-1 modulo 2^(P-s): a handful of (M, X) per e.
//
// Scaling x and d by powers of two changes nothing while no operation
// underflows or overflows; the kernels use the result only for
// 2^-900 <= |x| <= 2^900 and 2^-60 <= |d| <= 2^60, where that holds (an
// RN(x zl) in the subnormal range adds at most 2^-1075, far inside the margin
// of the wider e range).  Signs are symmetric.
// ---------------------------------------------------------------------------
namespace chdiv {

CUSTEN_CH_HD inline double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
CUSTEN_CH_HD inline float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// a * b as a 128-bit value hi:lo
CUSTEN_CH_HD inline void mulWide(unsigned long long a, unsigned long long b,
                                 unsigned long long& hi, unsigned long long& lo)
{
    const unsigned long long m32 = 0xffffffffull;
    const unsigned long long a0 = a & m32, a1 = a >> 32, b0 = b & m32, b1 = b >> 32;
    const unsigned long long p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const unsigned long long mid = (p00 >> 32) + (p01 & m32) + (p10 & m32);
    lo = (p00 & m32) | (mid << 32);
    hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// T = float (U = unsigned int, p = 24) or double (unsigned long long, 53).
// d must be normal.  True when q = RN(x zh + RN(x zl)) equals x / d for every
// listed x (and so for every x).
template <typename T, typename U, int p, int eMax = 16>
CUSTEN_CH_HD inline bool certified(T d)
{
    constexpr U mantMask = ((U)1 << (p - 1)) - 1;
    const U one = __builtin_bit_cast(U, (T)1);
    const U db = __builtin_bit_cast(U, d);
    const T dn = __builtin_bit_cast(T, (U)(one | (db & mantMask)));  // |d| scaled to [1, 2)
    const T zh = (T)1 / dn;
    const T zl = fmaT(-dn, zh, (T)1) / dn;
    const unsigned long long D = (unsigned long long)(db & mantMask) | (1ull << (p - 1));
    int s = 0;
    while (((D >> s) & 1ull) == 0) s++;
    // e = 2^s (odd) with |e| <= eMax needs 2^s <= eMax; for larger s no x
    // comes close to a midpoint at all
    if ((1ll << s) > eMax) return true;
    const unsigned long long Do = D >> s;
    unsigned long long inv = Do;  // Do^-1 modulo 2^64 (Newton, 3 -> 96 bits)
    for (int k = 0; k < 5; k++) inv *= 2ull - Do * inv;
    for (int P = p; P <= p + 1; P++) {
        const int Q = P - s;  // M is fixed modulo 2^Q
        const unsigned long long qMask = (1ull << Q) - 1;
        for (int e = -eMax; e <= eMax; e++) {
            if (e == 0 || (e % (1 << s)) != 0) continue;
            const long long eo = (long long)(e / (1 << s));
            if ((eo & 1) == 0) continue;
            const unsigned long long r = ((unsigned long long)(-eo) * inv) & qMask;
            const unsigned long long lo2 = 1ull << p, hi2 = 1ull << (p + 1);
            unsigned long long M = r;
            if (M < lo2) M += ((lo2 - M + qMask) >> Q) << Q;
            for (; M < hi2; M += 1ull << Q) {
                // X = (M D + e) / 2^P, exact by the choice of M
                unsigned long long h = 0, l = 0;
                mulWide(M, D, h, l);
                if (e > 0) {
                    const unsigned long long l2 = l + (unsigned long long)e;
                    h += l2 < l;
                    l = l2;
                } else {
                    const unsigned long long ue = (unsigned long long)(-e);
                    h -= l < ue;
                    l -= ue;
                }
                if ((l & ((1ull << P) - 1)) != 0) return false;  // cannot happen
                const unsigned long long X = (h << (64 - P)) | (l >> P);
                if (X < (1ull << (p - 1)) || X >= (1ull << p)) continue;
                const T x = __builtin_bit_cast(T, (U)(one | ((U)X & mantMask)));
                if (!(fmaT(x, zh, x * zl) == x / dn)) return false;
            }
        }
    }
    return true;
}

}  // namespace chdiv

// The divisor d of the column solve's constant run may use the two-operation
// division (q = fma(x, zh, x zl) with zh = RN(1/d), zl = RN(fma(-d, zh, 1) / d))
// for 2^-900 <= |x| <= 2^900.
CUSTEN_CH_HD inline bool chDivCertified(double d)
{
    const double ad = d < 0 ? -d : d;
    if (!(ad >= 0x1p-60 && ad <= 0x1p60)) return false;
    return chdiv::certified<double, unsigned long long, 53>(d);
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
