// const_runs_check.cpp -- the constant runs of the factored CH matrix
// (chConstantRuns, custen_amd/csrc/ch_setup.h) and the phased sweep of the
// fused column solve that uses them (custen_amd/csrc/ch_kernels.hip,
// colSolveKernel), restated on the CPU.
//
// For each grid size and coefficient it factors the matrix, checks that every
// row of each run equals the run's last row bitwise and that the row before
// the run does not, then solves random right-hand sides twice: with
// chPentSolve, and with the kernel's split of the sweeps into 16-row steps
// whose diagonals are read per row (outside the runs) or taken as the run's
// constants (inside).  Both use IEEE division; the phased sweep must give the
// same bits.  With "early" as the first argument the constant phase starts
// one step too soon, which must produce mismatches.
//
// usage: const_runs_check [early]   prints one line per case, then
// "cases <count> mismatches <count>"
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../custen_amd/csrc/ch_setup.h"

using custen::chConstantRuns;
using custen::chFactor;
using custen::chPentSolve;

namespace {

constexpr int SR = 16;

uint64_t state;
uint64_t next()  // SplitMix64
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t bits(double v)
{
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
}

// The column solve of colSolveKernel: forward rows 0 and 1, whole steps of SR
// rows (steps [0, jc) read the factors, steps [jc, nFull) use the constants
// of the forward run), the rows left one at a time; backward row m-2, whole
// steps (steps [0, jcb) in the backward run, then the rest from the factors),
// the rows left.
void phasedSolve(const double* fac, int m, int cfw, int cbw, int early, double* b)
{
    const double* ds = fac;
    const double* dl = ds + m;
    const double* dd = dl + m;
    const double* du = dd + m;
    const double* dw = du + m;
    const int nFull = (m - 2) / SR;
    const int kf = cfw < m - 1 ? cfw : m - 1, kb = cbw < m - 3 ? cbw : m - 3;
    const double k0 = ds[kf], k1 = dl[kf], k2 = dd[kf], ku = du[kb], kw = dw[kb];
    double b2 = b[0] / dd[0];
    b[0] = b2;
    double b1 = (b[1] - dl[1] * b2) / dd[1];
    b[1] = b1;
    int jc = (cfw - 2 + SR - 1) / SR;
    if (jc > nFull) jc = nFull;
    jc -= early;
    if (jc < 0) jc = 0;
    for (int j = 0; j < nFull; j++) {
        for (int k = 0; k < SR; k++) {
            const int i = 2 + j * SR + k;
            const bool c = j >= jc;
            const double t = b[i] - (c ? k0 : ds[i]) * b2;
            const double v = (t - (c ? k1 : dl[i]) * b1) / (c ? k2 : dd[i]);
            b[i] = v;
            b2 = b1;
            b1 = v;
        }
    }
    for (int i = 2 + nFull * SR; i < m; i++) {
        const double t = b[i] - ds[i] * b2;
        const double v = (t - dl[i] * b1) / dd[i];
        b[i] = v;
        b2 = b1;
        b1 = v;
    }
    double a1 = b1;
    double a0 = b2 - du[m - 2] * a1;
    b[m - 2] = a0;
    const int lowest0 = m - 3 - (SR - 1);
    int jcb = 0;
    if (lowest0 >= cbw) {
        jcb = (lowest0 - cbw) / SR + 1;
        if (jcb > nFull) jcb = nFull;
    }
    if (nFull > 0) jcb += early;
    if (jcb > nFull) jcb = nFull;
    for (int j = 0; j < nFull; j++) {
        for (int k = 0; k < SR; k++) {
            const int i = m - 3 - j * SR - k;
            const bool c = j < jcb;
            const double t = b[i] - (c ? ku : du[i]) * a0;
            const double v = t - (c ? kw : dw[i]) * a1;
            b[i] = v;
            a1 = a0;
            a0 = v;
        }
    }
    for (int i = m - 3 - nFull * SR; i >= 0; i--) {
        const double t = b[i] - du[i] * a0;
        const double v = t - dw[i] * a1;
        b[i] = v;
        a1 = a0;
        a0 = v;
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const int early = argc > 1 && std::strcmp(argv[1], "early") == 0 ? 1 : 0;
    const int sizes[] = {8, 9, 20, 21, 34, 50, 64, 100, 129, 256, 300, 512, 1024, 2048};
    // s = dt kappa / h^4 of the scheme: small steps to very stiff ones
    const double svals[] = {1e-4, 0.01, 0.37, 2.5, 40.0, 1e3, 1e6};
    int cases = 0;
    long mismatches = 0;
    state = 20240611;
    for (int n : sizes)
        for (double s : svals) {
            const int m = n - 2;
            std::vector<double> fac((size_t)7 * m);
            double omega[4];
            chFactor(fac.data(), m, s, -4 * s, 1 + 6 * s, -4 * s, s, omega);
            int fwd = 0, bwd = 0;
            chConstantRuns(fac.data(), m, fwd, bwd);
            const double* ds = fac.data();
            const double* dl = ds + m;
            const double* dd = dl + m;
            const double* du = dd + m;
            const double* dw = du + m;
            // the runs: rows [fwd, m-1] and [bwd, m-3] repeat their last row,
            // and the row before each run does not
            long bad = 0;
            for (int i = fwd; i < m; i++)
                bad += bits(ds[i]) != bits(ds[m - 1]) || bits(dl[i]) != bits(dl[m - 1]) ||
                       bits(dd[i]) != bits(dd[m - 1]);
            if (fwd > 0)
                bad += bits(ds[fwd - 1]) == bits(ds[m - 1]) &&
                       bits(dl[fwd - 1]) == bits(dl[m - 1]) && bits(dd[fwd - 1]) == bits(dd[m - 1]);
            for (int i = bwd; i <= m - 3; i++)
                bad += bits(du[i]) != bits(du[m - 3]) || bits(dw[i]) != bits(dw[m - 3]);
            if (bwd > 0)
                bad += bits(du[bwd - 1]) == bits(du[m - 3]) && bits(dw[bwd - 1]) == bits(dw[m - 3]);
            for (int rep = 0; rep < 4; rep++) {
                std::vector<double> x(m), y(m);
                for (int i = 0; i < m; i++) {
                    const double u = (double)(next() >> 11) * 0x1p-53;
                    x[i] = y[i] = 2 * u - 1;
                }
                chPentSolve(ds, dl, dd, du, dw, x.data(), m);
                phasedSolve(fac.data(), m, fwd, bwd, early, y.data());
                for (int i = 0; i < m; i++) bad += bits(x[i]) != bits(y[i]);
            }
            std::printf("n %d s %g fwd %d bwd %d mismatches %ld\n", n, s, fwd, bwd, bad);
            mismatches += bad;
            cases++;
        }
    std::printf("cases %d mismatches %ld\n", cases, mismatches);
    return 0;
}
