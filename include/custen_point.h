/*
 * custen_point.h -- point functions compiled into the caller's own stencil
 * kernel (extension; no counterpart in the reference).
 *
 * The reference applies a caller's point function through a device function
 * pointer (devArg1X / devArg1Y / devArg1XY, fetched with cudaMemcpyFromSymbol,
 * examples/src/2d_xy_p_fun.cu:177-178).  This library keeps that path
 * (cuStenCreate2D*Fun with `func`), but on gfx950 a call through a pointer
 * cannot be inlined: the callee reads the staged window and the coefficients
 * through generic pointers, one FLAT instruction per value, and the address
 * unit that serves them bounds the apply at about 4x the inlined time.
 *
 * A caller that compiles with hipcc can instead instantiate the apply kernel
 * for its function here, in its own translation unit, and attach the launcher
 * to the handle:
 *
 *     __device__ double nonLinRHS(double* d, double* c, int loc, int jump,
 *                                 int nx, int ny) { ... }
 *     cuStenCreate2DXYpFun(&h, ..., coe, func);          // as before
 *     custenSetPointLauncher(&h, custenPointLauncherXY<nonLinRHS>);
 *     cuStenCompute2DXYpFun(&h, HOST);                   // inlined from here on
 *
 * The function keeps the reference's signature and sees the same arguments:
 * `data` is the staged window in LDS, `coe` the coefficients in LDS, `loc` the
 * window centre (x, y) or top-left point (xy), `jump` the staged row length.
 * Inlined, its reads become LDS reads.  The results equal the pointer path's
 * bit for bit when the function is compiled with the same floating-point
 * contraction.  No relocatable device code is needed: the kernel lives in the
 * caller's code object and the library only calls the launcher.
 */
#ifndef CUSTEN_POINT_H
#define CUSTEN_POINT_H

#include "cuSten.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One region of an apply, as the library hands it to a launcher.  Output
 * row gy (0 <= gy < ny) is out + gy * ldOut; input rows are addressed through
 * the row map: rowWrap > 0 -> source row (gy mod rowWrap), otherwise
 * clamp(gy, rowMin, rowMax); columns wrap modulo nx when xWrap, else clamp.
 * Points with column in [wx0, wx1) and global row yGlobal0 + gy in
 * [wy0, wy1) are written; no other point is touched. */
typedef struct {
    const double* in;
    double* out;
    const double* coe;
    int numCoe;
    int dir;            /* 0 x, 1 y, 2 xy */
    long long ldIn, ldOut;
    int nx, ny;
    int H, L, V, T;     /* stencil extents: R = H-1-L, B = V-1-T */
    int xWrap;
    int rowWrap, rowMin, rowMax;
    int wx0, wx1, wy0, wy1, yGlobal0;
} custenPointArgs;

typedef hipError_t (*custenPointLauncher)(const custenPointArgs* args, hipStream_t stream);

/* Attach (or, with NULL, detach) a launcher to a point-function handle: every
 * later Compute of the handle runs it instead of calling the handle's
 * function pointer.  Records CUSTEN_ERROR_INVALID_VALUE for a handle that is
 * not a point-function handle. */
void custenSetPointLauncher(cuSten_t* pt_cuSten, custenPointLauncher launcher);

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus) && defined(__HIP__)
#include <hip/hip_runtime.h>

namespace custen_point {

constexpr int kLanes = 64;
constexpr int kStrip = 2 * kLanes;  /* columns per wave, two per lane */
constexpr int kRows = 8;            /* output rows per wave */

__device__ __forceinline__ int wrapIndex(int i, int n)
{
    if (i >= 0 && i < n) return i;
    const int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int sourceRow(const custenPointArgs& a, int gy)
{
    if (a.rowWrap > 0) return wrapIndex(gy, a.rowWrap);
    return gy < a.rowMin ? a.rowMin : (gy > a.rowMax ? a.rowMax : gy);
}

__device__ __forceinline__ int sourceCol(const custenPointArgs& a, int gx)
{
    if (a.xWrap) return wrapIndex(gx, a.nx);
    return gx < 0 ? 0 : (gx >= a.nx ? a.nx - 1 : gx);
}

/* LDS row length for a stencil H wide: 128 columns and the halo, even. */
__host__ __device__ constexpr int rowLength(int H) { return (kStrip + H + 1) & ~1; }

/* HC, VC: the stencil's width and height when known at compile time (0:
 * read from the arguments).  With them constant, the inlined function's
 * loops over its window unroll, its reads are issued together, and its
 * coefficient reads leave the loop over points. */
template <int DIR, int HC, int VC, class F>
__device__ __forceinline__ void applyBand(const custenPointArgs& a, F f)
{
    extern __shared__ double custenPointLds[];
    const int H = HC ? HC : a.H, V = VC ? VC : a.V, L = a.L;
    const int LD = rowLength(H);
    const int NR = kRows + V - 1;
    double* tile = custenPointLds;
    double* coeL = tile + NR * LD;
    const int lane = threadIdx.x;
    const int stripsX = (a.nx + kStrip - 1) / kStrip;
    const int x0 = (int)(blockIdx.x % stripsX) * kStrip;
    const int y0 = (int)(blockIdx.x / stripsX) * kRows;
    if (y0 >= a.ny) return;
    for (int k = lane; k < a.numCoe; k += kLanes) coeL[k] = a.coe[k];
    /* staged window: rows y0 - T .. y0 + kRows + B - 1, columns
     * x0 - L .. x0 + 127 + R (wrapped or clamped in the source address).
     * Lane i stages columns i, 64 + i and (the halo, lanes < H - 1) 128 + i
     * of each row; four rows' loads are issued before their stores. */
    const int c0 = sourceCol(a, x0 - L + lane);
    const int c1 = sourceCol(a, x0 - L + kLanes + lane);
    const bool halo = lane < H - 1;
    const int c2 = halo ? sourceCol(a, x0 - L + kStrip + lane) : c0;
    /* global-address-space view of the input: its loads cannot alias the
     * LDS stores */
    typedef __attribute__((address_space(1))) const double GlobalDouble;
    const GlobalDouble* in = (const GlobalDouble*)a.in;
    for (int r0 = 0; r0 < NR; r0 += 4) {
        double v[4][3];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int r = r0 + k < NR ? r0 + k : NR - 1;
            const GlobalDouble* src = in + (long long)sourceRow(a, y0 - a.T + r) * a.ldIn;
            v[k][0] = src[c0];
            v[k][1] = src[c1];
            v[k][2] = src[c2];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (r0 + k >= NR) break;
            double* row = tile + (r0 + k) * LD;
            row[lane] = v[k][0];
            row[kLanes + lane] = v[k][1];
            if (halo) row[kStrip + lane] = v[k][2];
        }
    }
    __syncthreads();
    int base = 2 * lane;
    if (DIR == 0) base += L;           /* window centre (V == 1) */
    else if (DIR == 1) base += a.T * LD;  /* window centre (H == 1) */
    const int rows = (a.ny - y0) < kRows ? (a.ny - y0) : kRows;
    const int gx = x0 + 2 * lane;
    double* orow = a.out + (long long)y0 * a.ldOut + gx;
    const bool both = gx + 1 < a.nx && gx >= a.wx0 && gx + 1 < a.wx1 &&
                      (((a.ldOut | x0) & 1) == 0) && (((unsigned long long)a.out & 15) == 0);
    for (int r = 0; r < rows; r++) {
        const int gyg = a.yGlobal0 + y0 + r;
        if (gyg < a.wy0 || gyg >= a.wy1) continue;
        const double v0 = f(tile, coeL, base + r * LD, LD, H, V);
        const double v1 = f(tile, coeL, base + r * LD + 1, LD, H, V);
        double* o = orow + (long long)r * a.ldOut;
        if (both) {
            /* both columns written and 16-byte aligned: one store */
            typedef double Pair __attribute__((ext_vector_type(2)));
            Pair p;
            p.x = v0;
            p.y = v1;
            *reinterpret_cast<Pair*>(o) = p;
        } else {
            if (gx < a.nx && gx >= a.wx0 && gx < a.wx1) o[0] = v0;
            if (gx + 1 < a.nx && gx + 1 >= a.wx0 && gx + 1 < a.wx1) o[1] = v1;
        }
    }
}

template <double (*FN)(double*, double*, int), int HC>
__global__ __launch_bounds__(kLanes) void kernelX(custenPointArgs a)
{
    applyBand<0, HC, 1>(a, [](double* d, double* c, int loc, int, int, int) {
        return FN(d, c, loc);
    });
}

template <double (*FN)(double*, double*, int, int), int VC>
__global__ __launch_bounds__(kLanes) void kernelY(custenPointArgs a)
{
    applyBand<1, 1, VC>(a, [](double* d, double* c, int loc, int jump, int, int) {
        return FN(d, c, loc, jump);
    });
}

template <double (*FN)(double*, double*, int, int, int, int), int HC, int VC>
__global__ __launch_bounds__(kLanes) void kernelXY(custenPointArgs a)
{
    applyBand<2, HC, VC>(a, [](double* d, double* c, int loc, int jump, int nx, int ny) {
        return FN(d, c, loc, jump, nx, ny);
    });
}

template <class K>
hipError_t launch(K kernel, const custenPointArgs* a, hipStream_t stream)
{
    if (a->nx <= 0 || a->ny <= 0) return hipSuccess;
    const long long strips = (a->nx + kStrip - 1) / kStrip;
    const long long blocks = strips * ((a->ny + kRows - 1) / kRows);
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * ((size_t)(kRows + a->V - 1) * rowLength(a->H) +
                                         (size_t)(a->numCoe > 0 ? a->numCoe : 1));
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kLanes), lds, stream, *a);
    return hipGetLastError();
}

}  // namespace custen_point

/* Launchers for a caller's devArg1X / devArg1Y / devArg1XY function.  The
 * common window sizes (3, 5, 7 and 9 points; 3 x 3 and 5 x 5) get kernels with
 * the size fixed at compile time; any other size runs the general kernel. */
template <double (*FN)(double*, double*, int)>
hipError_t custenPointLauncherX(const custenPointArgs* a, hipStream_t stream)
{
    using namespace custen_point;
    if (a->dir != 0) return hipErrorInvalidValue;
    switch (a->H) {
    case 3: return launch(kernelX<FN, 3>, a, stream);
    case 5: return launch(kernelX<FN, 5>, a, stream);
    case 7: return launch(kernelX<FN, 7>, a, stream);
    case 9: return launch(kernelX<FN, 9>, a, stream);
    default: return launch(kernelX<FN, 0>, a, stream);
    }
}
template <double (*FN)(double*, double*, int, int)>
hipError_t custenPointLauncherY(const custenPointArgs* a, hipStream_t stream)
{
    using namespace custen_point;
    if (a->dir != 1) return hipErrorInvalidValue;
    switch (a->V) {
    case 3: return launch(kernelY<FN, 3>, a, stream);
    case 5: return launch(kernelY<FN, 5>, a, stream);
    case 7: return launch(kernelY<FN, 7>, a, stream);
    case 9: return launch(kernelY<FN, 9>, a, stream);
    default: return launch(kernelY<FN, 0>, a, stream);
    }
}
template <double (*FN)(double*, double*, int, int, int, int)>
hipError_t custenPointLauncherXY(const custenPointArgs* a, hipStream_t stream)
{
    using namespace custen_point;
    if (a->dir != 2) return hipErrorInvalidValue;
    if (a->H == 3 && a->V == 3) return launch(kernelXY<FN, 3, 3>, a, stream);
    if (a->H == 5 && a->V == 5) return launch(kernelXY<FN, 5, 5>, a, stream);
    return launch(kernelXY<FN, 0, 0>, a, stream);
}
#endif

#endif /* CUSTEN_POINT_H */
