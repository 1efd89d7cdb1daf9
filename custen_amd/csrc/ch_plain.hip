// ch_plain.hip -- the reference-shaped Cahn-Hilliard step, kept as the plain
// per-grid baseline: the launch sequence of cuPentCahnADI/src/cuPentCahnADI.cu:528-596
// (findCBar, the two cuSten stencils, findRHS, transpose, pentSolveBatch,
// solveEnd, solveFull, transpose, pentSolveBatch, solveEnd, solveFull,
// findNew), each launch followed by a device synchronisation, thread-per-
// system solves over interleaved diagonals replicated per system
// (cuPentBatch.cu:119-198, BatchHyper.cu:195-259, 520-590).  The arithmetic is
// the fused engine's, so both paths produce identical bits.
#include <hip/hip_runtime.h>

#include "ch_plain.h"

#pragma clang fp contract(off)

namespace custen {
namespace plaindev {

constexpr int B2 = 32;  // 32 x 32 blocks for the elementwise kernels (BLOCK_X/BLOCK_Y)
constexpr int BI = 64;  // 64 threads for the solves (BLOCK_INV)

__global__ void findCBar(const double* cOld, const double* c, double* cBar, int n)
{
    const int x = blockIdx.x * B2 + threadIdx.x, y = blockIdx.y * B2 + threadIdx.y;
    if (x >= n || y >= n) return;
    const int64_t k = (int64_t)y * n + x;
    const double t = 2.0 * c[k];
    cBar[k] = t - cOld[k];
}

__global__ void findRHS(double* cOld, const double* c, double* cHalf, const double* nl, int n)
{
    const int x = blockIdx.x * B2 + threadIdx.x, y = blockIdx.y * B2 + threadIdx.y;
    if (x >= n || y >= n) return;
    const int64_t k = (int64_t)y * n + x;
    const double k23 = -(2.0 / 3.0);
    const double t = k23 * (c[k] - cOld[k]) + nl[k];
    cHalf[k] = cHalf[k] + t;
    cOld[k] = c[k];
}

__global__ void findNew(double* c, const double* cBar, const double* cHalf, int n)
{
    const int x = blockIdx.x * B2 + threadIdx.x, y = blockIdx.y * B2 + threadIdx.y;
    if (x >= n || y >= n) return;
    const int64_t k = (int64_t)y * n + x;
    c[k] = cBar[k] + cHalf[k];
}

__global__ void transpose(const double* in, double* out, int n)
{
    __shared__ double t[B2][B2 + 1];
    const int x = blockIdx.x * B2 + threadIdx.x, y = blockIdx.y * B2 + threadIdx.y;
    if (x < n && y < n) t[threadIdx.y][threadIdx.x] = in[(int64_t)y * n + x];
    __syncthreads();
    const int ox = blockIdx.y * B2 + threadIdx.x, oy = blockIdx.x * B2 + threadIdx.y;
    if (ox < n && oy < n) out[(int64_t)oy * n + ox] = t[threadIdx.x][threadIdx.y];
}

// b: m x batch interleaved (row-major, one system per column)
__global__ void pentSolveBatch(const double* ds, const double* dl, const double* d,
                               const double* du, const double* dw, double* b, int m, int batch)
{
    int rc = blockIdx.x * BI + threadIdx.x;
    if (rc >= batch) return;
    b[rc] = b[rc] / d[rc];
    int rp = rc;
    rc += batch;
    b[rc] = (b[rc] - dl[rc] * b[rp]) / d[rc];
    for (int i = 2; i < m; i++) {
        const int rpp = rc - batch;
        rp = rc;
        rc += batch;
        b[rc] = (b[rc] - ds[rc] * b[rpp] - dl[rc] * b[rp]) / d[rc];
    }
    int ra = rc;
    rc -= batch;
    b[rc] = b[rc] - du[rc] * b[ra];
    for (int i = m - 3; i >= 0; i--) {
        const int raa = rc + batch;
        ra = rc;
        rc -= batch;
        b[rc] = b[rc] - du[rc] * b[ra] - dw[rc] * b[raa];
    }
}

__global__ void solveEnd(double* data, double a, double b, double d, double e, double o11,
                         double o12, double o21, double o22, int nx, int nb)
{
    const int g = blockIdx.x * BI + threadIdx.x;
    if (g >= nb) return;
    const double nx2 = data[(int64_t)(nx - 2) * nb + g] -
                       (e * data[g] + a * data[(int64_t)(nx - 4) * nb + g] +
                        b * data[(int64_t)(nx - 3) * nb + g]);
    const double nx1 = data[(int64_t)(nx - 1) * nb + g] -
                       (d * data[g] + e * data[nb + g] + a * data[(int64_t)(nx - 3) * nb + g]);
    data[(int64_t)(nx - 2) * nb + g] = o11 * nx2 + o12 * nx1;
    data[(int64_t)(nx - 1) * nb + g] = o21 * nx2 + o22 * nx1;
}

__global__ void solveFull(double* data, const double* inv1, const double* inv2, int nx, int nb)
{
    const int x = blockIdx.x * B2 + threadIdx.x, y = blockIdx.y * B2 + threadIdx.y;
    if (x >= nb || y >= nx - 2) return;
    const double o2 = data[(int64_t)(nx - 2) * nb + x];
    const double o1 = data[(int64_t)(nx - 1) * nb + x];
    const int64_t k = (int64_t)y * nb + x;
    data[k] = data[k] - (inv1[k] * o2 + inv2[k] * o1);
}

}  // namespace plaindev

namespace {

hipError_t sync(hipStream_t st) { return hipStreamSynchronize(st); }

hipError_t cyclicInv(const PlainCH& p, double* data, hipStream_t st)
{
    const int n = p.n, m = n - 2;
    const int64_t mn = (int64_t)m * n;
    const double* M = p.multi;
    hipError_t e;
    hipLaunchKernelGGL(plaindev::pentSolveBatch, dim3((n + 63) / 64), dim3(64), 0, st, M, M + mn,
                       M + 2 * mn, M + 3 * mn, M + 4 * mn, data, m, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    hipLaunchKernelGGL(plaindev::solveEnd, dim3((n + 63) / 64), dim3(64), 0, st, data, p.a, p.b,
                       p.d, p.e, p.omega[0], p.omega[1], p.omega[2], p.omega[3], n, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    const dim3 grid((n + 31) / 32, (n + 31) / 32);
    hipLaunchKernelGGL(plaindev::solveFull, grid, dim3(32, 32), 0, st, data, M + 5 * mn,
                       M + 6 * mn, n, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace

hipError_t chPlainStep(const PlainCH& p, double* cCurr, double* cOld, hipStream_t st,
                       const std::function<hipError_t(double* out, const double* in, bool nonlinear)>& stencil)
{
    const int n = p.n;
    const dim3 grid((n + 31) / 32, (n + 31) / 32), blk(32, 32);
    hipError_t e;
    hipLaunchKernelGGL(plaindev::findCBar, grid, blk, 0, st, cOld, cCurr, p.cBar, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    if ((e = stencil(p.nonLin, cCurr, true)) != hipSuccess) return e;
    if ((e = stencil(p.cHalf, p.cBar, false)) != hipSuccess) return e;
    if ((e = sync(st)) != hipSuccess) return e;
    hipLaunchKernelGGL(plaindev::findRHS, grid, blk, 0, st, cOld, cCurr, p.cHalf, p.nonLin, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    hipLaunchKernelGGL(plaindev::transpose, grid, blk, 0, st, p.cHalf, cCurr, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    if ((e = cyclicInv(p, cCurr, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(plaindev::transpose, grid, blk, 0, st, cCurr, p.cHalf, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    if ((e = cyclicInv(p, p.cHalf, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(plaindev::findNew, grid, blk, 0, st, cCurr, p.cBar, p.cHalf, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = sync(st)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace custen
