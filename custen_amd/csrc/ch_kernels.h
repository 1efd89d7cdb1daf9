// ch_kernels.h -- Cahn-Hilliard ADI step kernels (batched over requests).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#if defined(__HIPCC__)
#define CUSTEN_HD __host__ __device__
#else
#define CUSTEN_HD
#endif

namespace custen {

// Per-request constants of the ADI scheme.  Everything a kernel needs for
// one resident simulation; an array of these lives in device memory.
struct CHDesc {
    double* f[2];       // the two time levels; at engine step s the current field
                        // is f[(s - admitStep) & 1] and the previous one the other;
                        // the final phase writes the new field over the previous one
    double* X;          // right-hand side / solution work, n*n
    const double* fac;  // factored diagonals + inverses: ds, dl, d, du, dw, inv1, inv2 (7 * m)
    double a, b, d, e;  // matrix diagonals (solveEnd)
    double omega[4];
    double wlin[25];    // 5x5 biharmonic RHS weights
    double wnl[9];      // 3x3 Laplacian of the nonlinear term
    int n;
    int admitStep;      // engine step at which f[0] held the current field
    int tileBase;       // first rhs/final tile of this request in the launch
    int rowBase;        // first x-solve row block
    int colBase;        // first y-solve column block
};

// Tile sizes of the work decomposition (host and device agree on these).
constexpr int kCHTileX = 64;       // rhs / final kernels: 64 x 16 points per workgroup
constexpr int kCHTileY = 16;
constexpr int kCHColBlock = 64;    // y-solve: columns per workgroup (one wave)
constexpr int kCHRowLds = 64 * 1024;  // x-solve: LDS bytes per row block
constexpr int kCHMaxN = 8190;         // one padded row (n + 1 doubles) must fit the LDS budget

CUSTEN_HD inline int chTiles(int n) { return ((n + kCHTileX - 1) / kCHTileX) * ((n + kCHTileY - 1) / kCHTileY); }
// Rows are staged in LDS with a stride of n + 1 doubles (odd stride: the
// threads of a row block walk their rows without bank conflicts).
CUSTEN_HD inline int chRowsPerBlock(int n)
{
    int r = kCHRowLds / ((n + 1) * 8);
    if (r > 64) r = 64;
    return r < 1 ? 1 : r;
}
CUSTEN_HD inline int chRowBlocks(int n) { int r = chRowsPerBlock(n); return (n + r - 1) / r; }
CUSTEN_HD inline int chColBlocks(int n) { return (n + kCHColBlock - 1) / kCHColBlock; }

// Launch the four phases of one step for `count` requests (descs in device
// memory; bases filled so that phase work items are contiguous per request).
hipError_t chLaunchRhs(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st);
hipError_t chLaunchXSolve(const CHDesc* descs, int count, int totalRowBlocks, hipStream_t st);
hipError_t chLaunchYSolve(const CHDesc* descs, int count, int totalColBlocks, hipStream_t st);
hipError_t chLaunchFinal(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st);

}  // namespace custen
