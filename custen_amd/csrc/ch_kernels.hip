// ch_kernels.hip -- the Cahn-Hilliard ADI step on gfx950, batched over any
// number of resident requests (each with its own grid size).
//
// One step of cuPentCahnADI (cuPentCahnADI/src/cuPentCahnADI.cu:528-596;
// serial form cuPentSpeedUp/serialCahnADITiming/serialCahnADI.c:1010-1047)
// in four phases:
//   rhs   : cBar = 2c - cOld, X = L5x5(cBar) + ((-2/3)(c - cOld) + L3x3(c^3 - c))
//   xsolve: cyclic pentadiagonal solve of every row of X (rows staged in LDS,
//           pentSolve + solveEnd + solveFull in place)
//   ysolve: cyclic solve of every column (thread per column; pentSolve +
//           solveEnd; the column solveFull is applied in `final`)
//   final : c_new = cBar + (X - (inv1 x~2 + inv2 x~1)), written over cOld
// The reference's two cublasDgeam transposes are not needed: rows are solved
// through LDS and columns with coalesced thread-per-column sweeps.
// Compiled with -ffp-contract=off: every operation is rounded as written, the
// stencil sums are explicit fma chains (the cuSten order).
#include <hip/hip_runtime.h>

#include "ch_kernels.h"

#pragma clang fp contract(off)

namespace custen {
namespace chdev {

__device__ __forceinline__ int findRequest(const CHDesc* d, int count, int item, int which)
{
    // largest r with base(r) <= item; bases ascend with r
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b = which == 0 ? d[mid].tileBase : (which == 1 ? d[mid].rowBase : d[mid].colBase);
        if (b <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int wrapN(int i, int n)
{
    if (i < 0) return i + n;
    if (i >= n) return i - n;
    return i;
}

// ---------------------------------------------------------------------------
// rhs
// ---------------------------------------------------------------------------
constexpr int RX = kCHTileX, RY = kCHTileY;
constexpr int CBW = RX + 4, CBH = RY + 4;  // cBar tile with a 2-point halo
constexpr int QW = RX + 2, QH = RY + 2;    // u^3 - u tile with a 1-point halo

__global__ __launch_bounds__(256) void rhsKernel(const CHDesc* __restrict__ descs, int count,
                                                 int step)
{
    __shared__ double cb[CBH * CBW];
    __shared__ double q[QH * QW];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 0);
    const CHDesc& D = descs[r];
    const int n = D.n;
    const int par = (step - D.admitStep) & 1;
    const double* __restrict__ c = D.f[par];
    const double* __restrict__ co = D.f[par ^ 1];
    const int local = item - D.tileBase;
    const int tilesX = (n + RX - 1) / RX;
    const int x0 = (local % tilesX) * RX;
    const int y0 = (local / tilesX) * RY;

    for (int idx = threadIdx.x; idx < CBH * CBW; idx += 256) {
        const int rr = idx / CBW, cc = idx - rr * CBW;
        int gy = y0 - 2 + rr, gx = x0 - 2 + cc;
        gy = ((gy % n) + n) % n;
        gx = ((gx % n) + n) % n;
        const int64_t k = (int64_t)gy * n + gx;
        const double cv = c[k];
        const double ov = co[k];
        const double t = 2.0 * cv;
        cb[idx] = t - ov;
        if (rr >= 1 && rr < CBH - 1 && cc >= 1 && cc < CBW - 1) {
            const double u2 = cv * cv;
            const double u3 = u2 * cv;
            q[(rr - 1) * QW + (cc - 1)] = u3 - cv;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = x0 + lane;
    const double k23 = -(2.0 / 3.0);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int ly = wave * 4 + rr;
        const int y = y0 + ly;
        if (x >= n || y >= n) continue;
        double lin = 0.0;
#pragma unroll
        for (int j = 0; j < 5; j++)
#pragma unroll
            for (int i = 0; i < 5; i++) lin = __builtin_fma(D.wlin[j * 5 + i], cb[(ly + j) * CBW + lane + i], lin);
        double nl = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) nl = __builtin_fma(D.wnl[j * 3 + i], q[(ly + j) * QW + lane + i], nl);
        const int64_t k = (int64_t)y * n + x;
        const double t1 = c[k] - co[k];
        const double t2 = k23 * t1;
        const double t3 = t2 + nl;
        D.X[k] = lin + t3;
    }
}

// ---------------------------------------------------------------------------
// x-direction solve: rows staged in LDS, one thread per row
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void xSolveKernel(const CHDesc* __restrict__ descs, int count)
{
    extern __shared__ double rows[];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 1);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int RB = chRowsPerBlock(n);
    const int stride = n + 1;
    const int row0 = (item - D.rowBase) * RB;
    const int nrows = min(RB, n - row0);
    double* __restrict__ X = D.X;

    for (int rr = 0; rr < nrows; rr++)
        for (int i = threadIdx.x; i < n; i += 256) rows[rr * stride + i] = X[(int64_t)(row0 + rr) * n + i];
    __syncthreads();

    if ((int)threadIdx.x < nrows) {
        double* b = rows + threadIdx.x * stride;
        const double* ds = D.fac;
        const double* dl = ds + m;
        const double* dd = dl + m;
        const double* du = dd + m;
        const double* dw = du + m;
        const double* inv1 = dw + m;
        const double* inv2 = inv1 + m;
        // pentSolve (serialCahnADI.c:161-241)
        double b2 = b[0] / dd[0];
        b[0] = b2;
        double b1 = (b[1] - dl[1] * b2) / dd[1];
        b[1] = b1;
        for (int i = 2; i < m; i++) {
            const double t = b[i] - ds[i] * b2;
            const double v = (t - dl[i] * b1) / dd[i];
            b[i] = v;
            b2 = b1;
            b1 = v;
        }
        // backward: b1 = b[m-1], b2 = b[m-2]
        double a1 = b[m - 1];
        double a0 = b[m - 2] - du[m - 2] * a1;
        b[m - 2] = a0;
        for (int i = m - 3; i >= 0; i--) {
            const double t = b[i] - du[i] * a0;
            const double v = t - dw[i] * a1;
            b[i] = v;
            a1 = a0;
            a0 = v;
        }
        // solveEnd (serialCahnADI.c:628-655)
        const double nx2 = b[n - 2] - (D.e * b[0] + D.a * b[n - 4] + D.b * b[n - 3]);
        const double nx1 = b[n - 1] - (D.d * b[0] + D.e * b[1] + D.a * b[n - 3]);
        const double x2 = D.omega[0] * nx2 + D.omega[1] * nx1;
        const double x1 = D.omega[2] * nx2 + D.omega[3] * nx1;
        b[n - 2] = x2;
        b[n - 1] = x1;
        // solveFull (serialCahnADI.c:661-678)
        for (int i = 0; i < m; i++) b[i] = b[i] - (inv1[i] * x2 + inv2[i] * x1);
    }
    __syncthreads();

    for (int rr = 0; rr < nrows; rr++)
        for (int i = threadIdx.x; i < n; i += 256) X[(int64_t)(row0 + rr) * n + i] = rows[rr * stride + i];
}

// ---------------------------------------------------------------------------
// y-direction solve: one thread per column (coalesced across the wave)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void ySolveKernel(const CHDesc* __restrict__ descs, int count)
{
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 2);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int col = (item - D.colBase) * kCHColBlock + (int)threadIdx.x;
    if (col >= n) return;
    double* __restrict__ X = D.X + col;
    const int64_t s = n;
    const double* ds = D.fac;
    const double* dl = ds + m;
    const double* dd = dl + m;
    const double* du = dd + m;
    const double* dw = du + m;

    double b2 = X[0] / dd[0];
    X[0] = b2;
    double b1 = (X[s] - dl[1] * b2) / dd[1];
    X[s] = b1;
#pragma unroll 4
    for (int i = 2; i < m; i++) {
        const double t = X[i * s] - ds[i] * b2;
        const double v = (t - dl[i] * b1) / dd[i];
        X[i * s] = v;
        b2 = b1;
        b1 = v;
    }
    double a1 = X[(m - 1) * s];
    double a0 = X[(m - 2) * s] - du[m - 2] * a1;
    X[(m - 2) * s] = a0;
#pragma unroll 4
    for (int i = m - 3; i >= 0; i--) {
        const double t = X[i * s] - du[i] * a0;
        const double v = t - dw[i] * a1;
        X[i * s] = v;
        a1 = a0;
        a0 = v;
    }
    const double x0v = a0;                // X[0]
    const double x1v = a1;                // X[1]
    const double xm2 = X[(n - 4) * s];
    const double xm1 = X[(n - 3) * s];
    const double nx2 = X[(n - 2) * s] - (D.e * x0v + D.a * xm2 + D.b * xm1);
    const double nx1 = X[(n - 1) * s] - (D.d * x0v + D.e * x1v + D.a * xm1);
    X[(n - 2) * s] = D.omega[0] * nx2 + D.omega[1] * nx1;
    X[(n - 1) * s] = D.omega[2] * nx2 + D.omega[3] * nx1;
}

// ---------------------------------------------------------------------------
// final: column solveFull + findNew, written over the previous field
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void finalKernel(const CHDesc* __restrict__ descs, int count,
                                                   int step)
{
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 0);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int par = (step - D.admitStep) & 1;
    const double* __restrict__ c = D.f[par];
    double* __restrict__ co = D.f[par ^ 1];
    const double* __restrict__ X = D.X;
    const double* inv1 = D.fac + 5 * m;
    const double* inv2 = inv1 + m;
    const int local = item - D.tileBase;
    const int tilesX = (n + RX - 1) / RX;
    const int x0 = (local % tilesX) * RX;
    const int y0 = (local / tilesX) * RY;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = x0 + lane;
    if (x >= n) return;
    const double e2 = X[(int64_t)(n - 2) * n + x];
    const double e1 = X[(int64_t)(n - 1) * n + x];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int y = y0 + wave * 4 + rr;
        if (y >= n) break;
        const int64_t k = (int64_t)y * n + x;
        double xf = X[k];
        if (y < m) xf = xf - (inv1[y] * e2 + inv2[y] * e1);
        const double t = 2.0 * c[k];
        const double cBar = t - co[k];
        co[k] = cBar + xf;
    }
}

}  // namespace chdev

hipError_t chLaunchRhs(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::rhsKernel, dim3(totalTiles), dim3(256), 0, st, descs, count, step);
    return hipGetLastError();
}

hipError_t chLaunchXSolve(const CHDesc* descs, int count, int totalRowBlocks, hipStream_t st)
{
    if (count <= 0 || totalRowBlocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::xSolveKernel, dim3(totalRowBlocks), dim3(256), kCHRowLds, st, descs,
                       count);
    return hipGetLastError();
}

hipError_t chLaunchYSolve(const CHDesc* descs, int count, int totalColBlocks, hipStream_t st)
{
    if (count <= 0 || totalColBlocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::ySolveKernel, dim3(totalColBlocks), dim3(kCHColBlock), 0, st, descs,
                       count);
    return hipGetLastError();
}

hipError_t chLaunchFinal(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::finalKernel, dim3(totalTiles), dim3(256), 0, st, descs, count, step);
    return hipGetLastError();
}

}  // namespace custen
