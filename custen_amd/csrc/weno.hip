// weno.hip -- fifth-order WENO upwind advection, periodic in x and y:
// out = u * Fx + v * Fy (cuSten/src/kernels/2d_xyADVWENO_p_kernel.cu:51-401).
//
// One workgroup of 256 threads covers a 64-column x 4*RPW-row tile.  The
// tile of the advected field plus a 3-point halo on every side is staged in
// LDS once; each lane then evaluates its column for RPW rows, reading the
// seven-point row and column crosses from LDS and u, v from HBM.  The upwind
// choice selects operands (no divergent branches): both signs run the same
// flux code on differently ordered differences.
//
// The flux is evaluated operation by operation in the reference's source
// order with contraction off; divisions are IEEE (correctly rounded), and
// the optional single-precision squares round to float, square in float and
// widen, as powf does on the reference's double arguments.  Results are
// therefore bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "stencil_tile.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

constexpr int kWenoTW = 64;            // columns per tile (one per lane)
constexpr int kWenoLD = kWenoTW + 6;   // LDS row: 3 halo columns either side

template <bool SS>
__device__ __forceinline__ double wenoSq(double d)
{
    if constexpr (SS) {
        const float f = static_cast<float>(d);
        const float f2 = f * f;
        return static_cast<double>(f2);
    } else {
        return d * d;
    }
}

// wenoSten (2d_xyADVWENO_p_kernel.cu:51-86), same operation order.
template <bool SS>
__device__ __forceinline__ double wenoFlux(double v1, double v2, double v3, double v4, double v5)
{
    const double phi1 = ((1.0 / 3.0) * v1 - (7.0 / 6.0) * v2) + (11.0 / 6.0) * v3;
    const double phi2 = (-(1.0 / 6.0) * v2 + (5.0 / 6.0) * v3) + (1.0 / 3.0) * v4;
    const double phi3 = ((1.0 / 3.0) * v3 + (5.0 / 6.0) * v4) - (1.0 / 6.0) * v5;
    const double s1 = (13.0 / 12.0) * wenoSq<SS>((v1 - 2.0 * v2) + v3) +
                      0.25 * wenoSq<SS>((v1 - 4.0 * v2) + 3.0 * v3);
    const double s2 = (13.0 / 12.0) * wenoSq<SS>((v2 - 2.0 * v3) + v4) + 0.25 * wenoSq<SS>(v2 - v4);
    const double s3 = (13.0 / 12.0) * wenoSq<SS>((v3 - 2.0 * v4) + v5) +
                      0.25 * wenoSq<SS>((3.0 * v3 - 4.0 * v4) + v5);
    const double a1 = 0.1 / wenoSq<SS>(s1 + 1e-06);
    const double a2 = 0.6 / wenoSq<SS>(s2 + 1e-06);
    const double a3 = 0.3 / wenoSq<SS>(s3 + 1e-06);
    const double denom = 1.0 / ((a1 + a2) + a3);
    const double w1 = a1 * denom, w2 = a2 * denom, w3 = a3 * denom;
    return (phi1 * w1 + phi2 * w2) + phi3 * w3;
}

// Five upwinded differences from d[1..6], d[k] = q[c + k - 3] - q[c + k - 4]
// scaled: u > 0 takes d[1..5], otherwise d[6..2] (2d_xyADVWENO_p_kernel.cu:288-346).
template <bool SS>
__device__ __forceinline__ double upwindFlux(const double (&d)[7], bool pos)
{
    return wenoFlux<SS>(pos ? d[1] : d[6], pos ? d[2] : d[5], pos ? d[3] : d[4],
                        pos ? d[4] : d[3], pos ? d[5] : d[2]);
}

template <int RPW, bool SS>
__global__ __launch_bounds__(kThreads) void wenoKernel(WenoLaunch a)
{
    constexpr int TH = kWaves * RPW;
    constexpr int NR = TH + 6;
    __shared__ double tile[NR * kWenoLD];

    const int tilesX = (a.nx + kWenoTW - 1) / kWenoTW;
    const int g = xcdRemap(blockIdx.x, gridDim.x);
    const int x0 = (g % tilesX) * kWenoTW;
    const int y0 = (g / tilesX) * TH;
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;

    // LDS column j holds global column x0 - 3 + j, LDS row r region row y0 - 3 + r
    for (int r = wave; r < NR; r += kWaves) {
        const int sy = mapRow(a.rows, y0 - 3 + r);
        const double* grow = a.in + (int64_t)sy * a.ldIn;
        double* lrow = tile + r * kWenoLD;
        lrow[3 + lane] = grow[wrapIndex(x0 + lane, a.nx)];
        if (lane < 6) {
            const int j = lane < 3 ? lane : kWenoTW + lane;
            lrow[j] = grow[wrapIndex(x0 - 3 + j, a.nx)];
        }
    }
    __syncthreads();

    const int x = x0 + lane;
    if (x >= a.nx) return;
#pragma unroll
    for (int r = 0; r < RPW; r++) {
        const int ly = wave * RPW + r;
        const int y = y0 + ly;
        if (y >= a.ny) break;
        const double* c = tile + (ly + 3) * kWenoLD + 3 + lane;
        double dx[7], dy[7];
#pragma unroll
        for (int k = 1; k < 7; k++) {
            dx[k] = (c[k - 3] - c[k - 4]) * a.coeDx;
            dy[k] = (c[(k - 3) * kWenoLD] - c[(k - 4) * kWenoLD]) * a.coeDy;
        }
        const int64_t is = (int64_t)y * a.ldSide + x;
        const double uu = a.u[is];
        const double vv = a.v[is];
        const double fx = upwindFlux<SS>(dx, uu > 0.0);
        const double fy = upwindFlux<SS>(dy, vv > 0.0);
        __builtin_nontemporal_store(uu * fx + vv * fy, a.out + (int64_t)y * a.ldOut + x);
    }
}

template <int RPW>
hipError_t launchWenoRpw(const WenoLaunch& w, hipStream_t st)
{
    constexpr int TH = kWaves * RPW;
    const int64_t tiles = (int64_t)((w.nx + kWenoTW - 1) / kWenoTW) * ((w.ny + TH - 1) / TH);
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    if (w.singleSquares)
        hipLaunchKernelGGL((wenoKernel<RPW, true>), dim3((unsigned)tiles), dim3(kThreads), 0, st, w);
    else
        hipLaunchKernelGGL((wenoKernel<RPW, false>), dim3((unsigned)tiles), dim3(kThreads), 0, st, w);
    return hipGetLastError();
}

}  // namespace dev

// 8 rows per wave unless 4 is asked for: measured faster at 8192^2 with both
// treatments of the squares (the kernel is VALU-bound; taller tiles re-read
// fewer halo rows and give each wave more independent points).
int wenoLdsBytes(int rowsPerWave)
{
    const int rpw = rowsPerWave == 4 ? 4 : 8;
    return (dev::kWaves * rpw + 6) * dev::kWenoLD * (int)sizeof(double);
}

hipError_t launchWeno(const WenoLaunch& w, hipStream_t st)
{
    if (w.nx < 1 || w.ny < 1) return hipSuccess;
    if (w.rowsPerWave == 4) return dev::launchWenoRpw<4>(w, st);
    return dev::launchWenoRpw<8>(w, st);
}

}  // namespace custen
