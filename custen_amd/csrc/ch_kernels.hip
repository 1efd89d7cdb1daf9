// ch_kernels.hip -- the Cahn-Hilliard ADI step on gfx950, batched over any
// number of resident requests (each with its own grid size).
//
// One step of cuPentCahnADI (cuPentCahnADI/src/cuPentCahnADI.cu:528-596;
// serial form cuPentSpeedUp/serialCahnADITiming/serialCahnADI.c:1010-1047)
// in four phases:
//   rhs   : cBar = 2c - cOld, X = L5x5(cBar) + ((-2/3)(c - cOld) + L3x3(c^3 - c))
//   xsolve: Y = X^T, cyclic pentadiagonal solve of every column of Y
//           (pentSolve + solveEnd), X = Y^T with solveFull applied on the way
//   ysolve: cyclic solve of every column of X (pentSolve + solveEnd)
//   final : c_new = cBar + (X - (inv1 x~2 + inv2 x~1)), written over cOld
// Every solve is a thread-per-column sweep (coalesced across the wave) with a
// block of row loads in flight ahead of the recurrence; the transposes
// replace the reference's cublasDgeam calls (cuPentCahnADI.cu:552, 566).
// Compiled with -ffp-contract=off: every operation is rounded as written, the
// stencil sums are explicit fma chains (the cuSten order).
#include <hip/hip_runtime.h>

#include "ch_kernels.h"

#pragma clang fp contract(off)

namespace custen {
namespace chdev {

// Global-address-space views of the descriptors' field pointers: accesses
// through them are global (not flat) instructions, which the wait counters
// track separately from LDS traffic.
using gdouble = __attribute__((address_space(1))) double;

__device__ __forceinline__ int findRequest(const CHDesc* d, int count, int item, int which)
{
    // largest r with base(r) <= item; bases ascend with r
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b = which == 0 ? d[mid].tileBase : (which == 2 ? d[mid].colBase : d[mid].trBase);
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
    const gdouble* __restrict__ c = (const gdouble*)D.f[par];
    const gdouble* __restrict__ co = (const gdouble*)D.f[par ^ 1];
    gdouble* __restrict__ Xo = (gdouble*)D.X;
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
        Xo[k] = lin + t3;
    }
}

// ---------------------------------------------------------------------------
// transpose (64 x 64 tiles through LDS).  corr = 0: Y = X^T.  corr = 1:
// X = Y'^T where Y' is Y with the column solveFull of the x-direction solves
// applied (serialCahnADI.c:661-678): y_j -= inv1[j] y~2 + inv2[j] y~1, j < n-2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void transposeKernel(const CHDesc* __restrict__ descs, int count,
                                                       int corr)
{
    __shared__ double t[kCHTr][kCHTr + 1];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 3);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int tiles = (n + kCHTr - 1) / kCHTr;
    const int local = item - D.trBase;
    const int R0 = (local / tiles) * kCHTr;  // source rows
    const int C0 = (local % tiles) * kCHTr;  // source columns
    const gdouble* __restrict__ src = (const gdouble*)(corr ? D.Y : D.X);
    gdouble* __restrict__ dst = (gdouble*)(corr ? D.X : D.Y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = C0 + lane;
    double e2 = 0.0, e1 = 0.0;
    const gdouble* inv1 = (const gdouble*)D.fac + 5 * m;
    const gdouble* inv2 = inv1 + m;
    if (corr && c < n) {
        e2 = src[(int64_t)(n - 2) * n + c];
        e1 = src[(int64_t)(n - 1) * n + c];
    }
#pragma unroll 4
    for (int k = wave; k < kCHTr; k += 4) {
        const int row = R0 + k;
        if (row < n && c < n) {
            double v = src[(int64_t)row * n + c];
            if (corr && row < m) v = v - (inv1[row] * e2 + inv2[row] * e1);
            t[k][lane] = v;
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = wave; k < kCHTr; k += 4) {
        const int row = C0 + k;  // destination row = source column
        const int col = R0 + lane;
        if (row < n && col < n) dst[(int64_t)row * n + col] = t[lane][k];
    }
}

// ---------------------------------------------------------------------------
// column solve: one thread per column (coalesced across the wave), rows
// streamed with a block of loads in flight ahead of the recurrence.
// pentSolve (serialCahnADI.c:161-241) + solveEnd (:628-655); the column's
// solveFull is applied by the next phase.  which = 0: Y, which = 1: X.
// ---------------------------------------------------------------------------
constexpr int PF = 16;  // rows of the column in registers ahead of the recurrence
constexpr int DB = 64;  // rows of the shared diagonals staged in LDS per block

// One PF-row block of the forward sweep (rows i0 .. i0+PF-1, guarded against
// m when !FULL): diagonals from LDS into registers first, then the recurrence
// x_i = ((b_i - ds_i x_{i-2}) - dl_i x_{i-1}) / d_i.
template <bool FULL>
__device__ __forceinline__ void forwardBlock(gdouble* __restrict__ base, int64_t s, int col, int i0,
                                             int m, const double (*dg)[DB], int o,
                                             const double (&cur)[PF], double& b2, double& b1)
{
    double a[PF], l[PF], d[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        a[k] = dg[0][o + k];
        l[k] = dg[1][o + k];
        d[k] = dg[2][o + k];
    }
#pragma unroll
    for (int k = 0; k < PF; k++) {
        if (FULL || i0 + k < m) {
            const double t = cur[k] - a[k] * b2;
            const double v = (t - l[k] * b1) / d[k];
            base[(int64_t)(i0 + k) * s + col] = v;
            b2 = b1;
            b1 = v;
        }
    }
}

// One PF-row block of the backward sweep (rows i0, i0-1, ..., i0-PF+1):
// x_i = (y_i - du_i x_{i+1}) - dw_i x_{i+2}.
template <bool FULL>
__device__ __forceinline__ void backwardBlock(gdouble* __restrict__ base, int64_t s, int col, int i0,
                                              const double (*dg)[DB], int o,
                                              const double (&cur)[PF], double& a0, double& a1)
{
    double u[PF], w[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        u[k] = dg[0][o + k];
        w[k] = dg[1][o + k];
    }
#pragma unroll
    for (int k = 0; k < PF; k++) {
        if (FULL || i0 - k >= 0) {
            const double t = cur[k] - u[k] * a0;
            const double v = t - w[k] * a1;
            base[(int64_t)(i0 - k) * s + col] = v;
            a1 = a0;
            a0 = v;
        }
    }
}

__global__ __launch_bounds__(64) void colSolveKernel(const CHDesc* __restrict__ descs, int count,
                                                     int which)
{
    // Diagonals of the current and the next DB-row block (every column of a
    // request shares one matrix): staged in LDS one block ahead, read as
    // broadcasts.  Forward uses ds, dl, d; backward du, dw.
    __shared__ double sd[2][3][DB];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 2);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int lane = (int)threadIdx.x;
    // Lanes past the last column repeat the last column: they read the same
    // values and store the same bits, which keeps every access unconditional.
    const int colRaw = (item - D.colBase) * kCHColBlock + lane;
    const bool live = colRaw < n;
    const int col = live ? colRaw : n - 1;
    gdouble* __restrict__ base = (gdouble*)(which ? D.X : D.Y);
    const int64_t s = n;
    const gdouble* __restrict__ ds = (const gdouble*)D.fac;
    const gdouble* __restrict__ dl = ds + m;
    const gdouble* __restrict__ dd = dl + m;
    const gdouble* __restrict__ du = dd + m;
    const gdouble* __restrict__ dw = du + m;

    // ---- forward substitution: rows 0, 1, then blocks of DB rows from row 2
    double b2 = base[col] / dd[0];
    base[col] = b2;
    double b1 = (base[s + col] - dl[1] * b2) / dd[1];
    base[s + col] = b1;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;  // lane's rows of the next diagonal block
    {
        const int rr = 2 + lane;
        if (rr < m) {
            sd[0][0][lane] = ds[rr];
            sd[0][1][lane] = dl[rr];
            sd[0][2][lane] = dd[rr];
        }
        const int rn = 2 + DB + lane;
        if (rn < m) {
            p0 = ds[rn];
            p1 = dl[rn];
            p2 = dd[rn];
        }
    }
    __syncthreads();
    double cur[PF], nxt[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) cur[k] = (2 + k < m) ? base[(int64_t)(2 + k) * s + col] : 0.0;
    for (int i0 = 2; i0 < m; i0 += PF) {
        const int rel = i0 - 2;
        const int blk = rel / DB;
        if (rel > 0 && rel % DB == 0) {
            // block blk starts: its diagonals (prefetched during block blk-1)
            // go to LDS buffer blk & 1, then block blk+1 is prefetched
            sd[blk & 1][0][lane] = p0;
            sd[blk & 1][1][lane] = p1;
            sd[blk & 1][2][lane] = p2;
            const int rn = 2 + (blk + 1) * DB + lane;
            if (rn < m) {
                p0 = ds[rn];
                p1 = dl[rn];
                p2 = dd[rn];
            }
            __syncthreads();
        }
        if (i0 + 2 * PF <= m) {
#pragma unroll
            for (int k = 0; k < PF; k++) nxt[k] = base[(int64_t)(i0 + PF + k) * s + col];
        } else {
#pragma unroll
            for (int k = 0; k < PF; k++) {
                const int i = i0 + PF + k;
                nxt[k] = i < m ? base[(int64_t)i * s + col] : 0.0;
            }
        }
        const double(*dg)[DB] = sd[blk & 1];
        const int o = rel - blk * DB;
        if (i0 + PF <= m) forwardBlock<true>(base, s, col, i0, m, dg, o, cur, b2, b1);
        else forwardBlock<false>(base, s, col, i0, m, dg, o, cur, b2, b1);
#pragma unroll
        for (int k = 0; k < PF; k++) cur[k] = nxt[k];
    }

    // ---- backward substitution: b1 = x[m-1], b2 = b[m-2]; blocks of DB rows
    // descending from row m-3
    const double xm1 = b1;
    double a1 = b1;
    double a0 = b2 - du[m - 2] * a1;
    base[(int64_t)(m - 2) * s + col] = a0;
    const double xm2 = a0;
    __syncthreads();  // every read of the forward blocks is done
    {
        const int rr = m - 3 - lane;
        if (rr >= 0) {
            sd[0][0][lane] = du[rr];
            sd[0][1][lane] = dw[rr];
        }
        const int rn = m - 3 - DB - lane;
        p0 = p1 = 0.0;
        if (rn >= 0) {
            p0 = du[rn];
            p1 = dw[rn];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PF; k++) cur[k] = (m - 3 - k >= 0) ? base[(int64_t)(m - 3 - k) * s + col] : 0.0;
    for (int i0 = m - 3; i0 >= 0; i0 -= PF) {
        const int rel = m - 3 - i0;
        const int blk = rel / DB;
        if (rel > 0 && rel % DB == 0) {
            sd[blk & 1][0][lane] = p0;
            sd[blk & 1][1][lane] = p1;
            const int rn = m - 3 - (blk + 1) * DB - lane;
            if (rn >= 0) {
                p0 = du[rn];
                p1 = dw[rn];
            }
            __syncthreads();
        }
        if (i0 - 2 * PF + 1 >= 0) {
#pragma unroll
            for (int k = 0; k < PF; k++) nxt[k] = base[(int64_t)(i0 - PF - k) * s + col];
        } else {
#pragma unroll
            for (int k = 0; k < PF; k++) {
                const int i = i0 - PF - k;
                nxt[k] = i >= 0 ? base[(int64_t)i * s + col] : 0.0;
            }
        }
        const double(*dg)[DB] = sd[blk & 1];
        const int o = rel - blk * DB;
        if (i0 - PF + 1 >= 0) backwardBlock<true>(base, s, col, i0, dg, o, cur, a0, a1);
        else backwardBlock<false>(base, s, col, i0, dg, o, cur, a0, a1);
#pragma unroll
        for (int k = 0; k < PF; k++) cur[k] = nxt[k];
    }
    if (!live) return;
    // solveEnd: a0 = x[0], a1 = x[1]
    gdouble* __restrict__ Xc = base + col;
    const double nx2 = Xc[(n - 2) * s] - (D.e * a0 + D.a * xm2 + D.b * xm1);
    const double nx1 = Xc[(n - 1) * s] - (D.d * a0 + D.e * a1 + D.a * xm1);
    Xc[(n - 2) * s] = D.omega[0] * nx2 + D.omega[1] * nx1;
    Xc[(n - 1) * s] = D.omega[2] * nx2 + D.omega[3] * nx1;
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
    const gdouble* __restrict__ c = (const gdouble*)D.f[par];
    gdouble* __restrict__ co = (gdouble*)D.f[par ^ 1];
    const gdouble* __restrict__ X = (const gdouble*)D.X;
    const gdouble* inv1 = (const gdouble*)D.fac + 5 * m;
    const gdouble* inv2 = inv1 + m;
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

hipError_t chLaunchTranspose(const CHDesc* descs, int count, int totalTiles, int corr,
                             hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::transposeKernel, dim3(totalTiles), dim3(256), 0, st, descs, count,
                       corr);
    return hipGetLastError();
}

hipError_t chLaunchColSolve(const CHDesc* descs, int count, int totalColBlocks, int which,
                            hipStream_t st)
{
    if (count <= 0 || totalColBlocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::colSolveKernel, dim3(totalColBlocks), dim3(kCHColBlock), 0, st,
                       descs, count, which);
    return hipGetLastError();
}

hipError_t chLaunchFinal(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::finalKernel, dim3(totalTiles), dim3(256), 0, st, descs, count, step);
    return hipGetLastError();
}

}  // namespace custen
