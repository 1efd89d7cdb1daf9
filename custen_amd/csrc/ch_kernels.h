// ch_kernels.h -- Cahn-Hilliard ADI step kernels (batched over requests).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>

#include "../../include/custen_ch.h"

#if defined(__HIPCC__)
#define CUSTEN_HD __host__ __device__
#else
#define CUSTEN_HD
#endif

namespace custen {

// Per-request constants of the ADI scheme.  Everything a kernel needs for
// one resident simulation; an array of these lives in device memory.
struct CHDesc {
    double* f[3];       // field buffers used in turn (the first `ring`): at engine
                        // step s, with t = s - admitStep, the current field is
                        // f[t % ring], the previous one f[(t + ring - 1) % ring],
                        // and the step writes the new one to f[(t + 1) % ring]
                        // (ring 2: over the previous field, in place)
    double* X;          // solution work, n*n (y-direction solves)
    double* Y;          // transposed right-hand side, n*n (x-direction solves)
    const double* fac;  // factored diagonals + inverses: ds, dl, d, du, dw, inv1, inv2 (7 * m)
    double a, b, d, e;  // matrix diagonals (solveEnd)
    double omega[4];
    double wlin[25];    // 5x5 biharmonic RHS weights
    double wnl[9];      // 3x3 Laplacian of the nonlinear term
    int n;
    int ring;           // 2 or 3 field buffers; finalRhs needs 3
    int admitStep;      // engine step at which f[0] held the current field
    int tileBase;       // first rhs/final tile of this request in the launch
    int colBase;        // first column-solve block
    int trBase;         // first transpose tile
    int constFwd1;      // 1 + first row of the forward factors' constant run
    int constBwd1;      // 1 + first row of the backward factors' constant run
                        // (chConstantRuns); 0 = not known: every row is read
                        // from the factors
    int fastDiv;        // 1: the forward run's pivot passes chDivCertified, and
                        // its steps divide in two operations
};

// Tile sizes of the work decomposition (host and device agree on these).
constexpr int kCHTileX = 64;       // rhs / final kernels: 64 x 16 points per workgroup
constexpr int kCHTileY = 16;
constexpr int kCHColBlock = 64;    // column solves: columns per workgroup (one wave)
constexpr int kCHTr = 64;          // transpose tile edge
constexpr int kCHMaxN = 32768;

CUSTEN_HD inline int chTiles(int n) { return ((n + kCHTileX - 1) / kCHTileX) * ((n + kCHTileY - 1) / kCHTileY); }
CUSTEN_HD inline int chColBlocks(int n) { return (n + kCHColBlock - 1) / kCHColBlock; }
CUSTEN_HD inline int chTrTiles(int n) { const int t = (n + kCHTr - 1) / kCHTr; return t * t; }

// The phases of one step for `count` requests (descs in device memory; the
// bases are filled so that each phase's work items are contiguous per
// request):
//   rhs        Y = RHS(c, cOld)^T (the right-hand side, written transposed)
//   colSolve   cyclic solve of every column of Y          (which = 0)
//   transpose  X = (Y with the column solveFull applied)^T
//   colSolve   cyclic solve of every column of X          (which = 1)
//   final      c_new = cBar + (X with the column solveFull applied)
//   finalRhs   final of step `step` and rhs of step + 1 in one pass: c_new is
//              computed over the tile and its halo from X, c and cOld, written
//              to the third field buffer (ring 3 only), and the next
//              right-hand side is formed from it and c
hipError_t chLaunchRhs(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st);
hipError_t chLaunchTranspose(const CHDesc* descs, int count, int totalTiles, hipStream_t st);
hipError_t chLaunchColSolve(const CHDesc* descs, int count, int totalColBlocks, int which,
                            hipStream_t st);
hipError_t chLaunchFinal(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st);
hipError_t chLaunchFinalRhs(const CHDesc* descs, int count, int totalTiles, int step,
                            hipStream_t st);

// Validates `p` and fills the constants of a descriptor: a, b, d, e, the two
// weight stencils and n (pointers, omega and the work bases are left zero);
// cDiag is the main diagonal 1 + 6 sigma.  False with `why` when invalid.
bool chDescribe(const custenCHParams& p, CHDesc& d, double& cDiag, std::string& why);

}  // namespace custen
