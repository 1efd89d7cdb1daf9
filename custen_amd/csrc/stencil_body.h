// stencil_body.h -- the register-blocked strip computation shared by the
// single-grid kernel (stencil_kernels.hip) and the batched-grids kernel
// (batch_kernels.hip).
#pragma once

#include "stencil_tile.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

// Process column strip `g` of the region described by `a`: all of its row
// chunks.  `w` holds the H*V weights, `tile` the LDS staging buffer of
// (4*RPW + V - 1) x ldsRowLength(H) elements.
template <typename T, int H, int V, int RPW, Xform X>
__device__ __forceinline__ void stencilStrip(const StencilLaunch& a, const T (&w)[H * V],
                                             T* __restrict__ tile, int g)
{
    constexpr int TH = kWaves * RPW;
    constexpr int NR = TH + V - 1;
    constexpr int LD = ldsRowLength(H);

    const T* __restrict__ in = static_cast<const T*>(a.in);
    T* __restrict__ out = static_cast<T*>(a.out);

    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int L = a.L;
    const int R = H - 1 - L;
    const int LP = (L + 1) & ~1;
    const int nx = a.nx;
    const int tilesX = (nx + kTW - 1) / kTW;
    const int tx = g % tilesX;
    const int tyg = g / tilesX;
    const int x0 = tx * kTW;
    const bool inside = x0 + kTW <= nx;
    const bool vecIn = a.vecIn && inside;
    const bool fullRow = a.vecOut && inside && x0 >= a.wx0 && x0 + kTW <= a.wx1;

    for (int c = 0; c < a.chunks; c++) {
        const int y0 = (tyg * a.chunks + c) * TH;
        if (y0 >= a.ny) break;
        if (c > 0) __syncthreads();
        stageTile<T, X>(tile, LD, NR, in, a.ldIn, a.rows, y0, a.T, x0, nx, L, R, LP, a.xWrap,
                        vecIn, wave, lane);
        __syncthreads();

        T acc[RPW][2];
#pragma unroll
        for (int r = 0; r < RPW; r++) acc[r][0] = acc[r][1] = T(0);

        const T* base = tile + (wave * RPW) * LD + (LP - L) + 2 * lane;
#pragma unroll
        for (int ir = 0; ir < RPW + V - 1; ir++) {
            T v[H + 1];
#pragma unroll
            for (int k = 0; k <= H; k++) v[k] = base[ir * LD + k];
#pragma unroll
            for (int r = 0; r < RPW; r++) {
                const int j = ir - r;
                if (j >= 0 && j < V) {
#pragma unroll
                    for (int i = 0; i < H; i++) {
                        acc[r][0] = fmaT(w[j * H + i], v[i], acc[r][0]);
                        acc[r][1] = fmaT(w[j * H + i], v[i + 1], acc[r][1]);
                    }
                }
            }
        }

#pragma unroll
        for (int r = 0; r < RPW; r++) {
            const int gy = y0 + wave * RPW + r;
            if (gy >= a.ny) break;
            const int gyg = a.yGlobal0 + gy;
            if (gyg < a.wy0 || gyg >= a.wy1) continue;
            storePair<T>(out + (int64_t)gy * a.ldOut, x0 + 2 * lane, acc[r][0], acc[r][1], fullRow,
                         nx, a.wx0, a.wx1, a.zeroRight, R);
        }
    }
}

}  // namespace dev
}  // namespace custen
