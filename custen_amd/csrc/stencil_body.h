// stencil_body.h -- the register-blocked strip computation shared by the
// single-grid kernel (stencil_kernels.hip) and the batched-grids kernel
// (batch.hip).
#pragma once

#include "stencil_tile.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

// Process strip (tx, tyg) of the region described by `a`: column strip tx
// (64 * CPL columns), row chunks [tyg * chunks, (tyg + 1) * chunks).  `w` holds
// the H*V weights, `tile` the LDS staging buffer of
// (4*RPW + V - 1) x ldsRowLengthW(64 * CPL, H) elements.
template <typename T, int H, int V, int RPW, Xform X, bool FAST, int CPL = 2>
__device__ __forceinline__ void stencilStripImpl(const StencilLaunch& a, const T (&w)[H * V],
                                                 T* __restrict__ tile, int tx, int tyg)
{
    constexpr int TH = kWaves * RPW;
    constexpr int NR = TH + V - 1;
    constexpr int TW = kLanes * CPL;
    constexpr int LD = ldsRowLengthW(TW, H);

    const Glob<T>* __restrict__ in = globalPtr<T>(a.in);
    Glob<T>* __restrict__ out = globalPtr<T>(a.out);

    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int L = a.L;
    const int R = H - 1 - L;
    const int LP = (L + 1) & ~1;
    const int nx = a.nx;
    const int x0 = tx * TW;
    const bool inside = x0 + TW <= nx;
    const bool vecIn = (CPL == 1 || a.vecIn) && inside;
    const bool fullRow = (CPL == 1 || a.vecOut) && inside && x0 >= a.wx0 && x0 + TW <= a.wx1;

    const int y00 = tyg * a.chunks * TH;
    if (y00 >= a.ny) return;
    const int B = V - 1 - a.T;
    // Chunk c needs region rows [y0 - T, y0 + TH + B).  The first chunk loads
    // all NR of them; every later chunk keeps its first V-1 rows from the
    // previous tile (LDS rows [TH, NR) move to [0, V-1)) and loads only the
    // TH new rows [y0 + B, y0 + TH + B), prefetched while the previous chunk
    // computes.
    {
        StagedRows<T, NR, CPL> first;
        loadRows<T, NR, FAST, CPL>(first, in, a.ldIn, a.rows, y00, a.T, x0, nx, L, R, a.xWrap,
                                   vecIn, wave, lane);
        storeRows<T, X, NR, CPL>(first, tile, LD, L, R, LP, wave, lane);
    }
    constexpr int NC = ((V - 1) * LD + kThreads - 1) / kThreads;  // carried values per thread
    StagedRows<T, TH, CPL> st;
    for (int c = 0; c < a.chunks; c++) {
        const int y0 = y00 + c * TH;
        if (c > 0) {
            T carry[NC > 0 ? NC : 1];
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const int idx = threadIdx.x + k * kThreads;
                if (idx < (V - 1) * LD) carry[k] = tile[TH * LD + idx];
            }
            __syncthreads();  // every wave is done with the previous chunk
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const int idx = threadIdx.x + k * kThreads;
                if (idx < (V - 1) * LD) tile[idx] = carry[k];
            }
            storeRows<T, X, TH, CPL>(st, tile + (V - 1) * LD, LD, L, R, LP, wave, lane);
        }
        __syncthreads();
        const bool more = (c + 1 < a.chunks) && (y0 + TH < a.ny);
        // prefetch the next chunk's new rows; they stay in flight during the compute
        if (more)
            loadRows<T, TH, FAST, CPL>(st, in, a.ldIn, a.rows, y0 + TH + B, 0, x0, nx, L, R,
                                       a.xWrap, vecIn, wave, lane);

        T acc[RPW][CPL];
#pragma unroll
        for (int r = 0; r < RPW; r++)
#pragma unroll
            for (int e = 0; e < CPL; e++) acc[r][e] = T(0);

        const T* base = tile + (wave * RPW) * LD + (LP - L) + CPL * lane;
#pragma unroll
        for (int ir = 0; ir < RPW + V - 1; ir++) {
            T v[H + CPL - 1];
#pragma unroll
            for (int k = 0; k < H + CPL - 1; k++) v[k] = base[ir * LD + k];
#pragma unroll
            for (int r = 0; r < RPW; r++) {
                const int j = ir - r;
                if (j >= 0 && j < V) {
#pragma unroll
                    for (int i = 0; i < H; i++)
#pragma unroll
                        for (int e = 0; e < CPL; e++)
                            acc[r][e] = fmaT(w[j * H + i], v[i + e], acc[r][e]);
                }
            }
        }

#pragma unroll
        for (int r = 0; r < RPW; r++) {
            const int gy = y0 + wave * RPW + r;
            if (gy >= a.ny) break;
            const int gyg = a.yGlobal0 + gy;
            if (gyg < a.wy0 || gyg >= a.wy1) continue;
            storeLane<T, CPL>(out + (int64_t)gy * a.ldOut, x0 + CPL * lane, acc[r], fullRow, nx,
                              a.wx0, a.wx1, a.zeroRight, R);
        }
        if (!more) break;
    }
}

// True when column strip tx is a full 128 columns of pair-aligned rows and
// one wrap suffices for row indices: such strips run the FAST instantiation
// (only a ragged last strip, or unaligned data, takes the general path).
__host__ __device__ inline bool stripIsInterior(const StencilLaunch& a, int tx, int cpl = 2)
{
    const int tw = kLanes * cpl;
    const int x0 = tx * tw;
    return (cpl == 1 || (a.vecIn && a.vecOut)) && x0 + tw <= a.nx &&
           (a.rows.wrap ? a.rows.rowWrap >= a.V : true);
}

// Strip g of a single-grid launch -> (tx, tyg): the listed edge strips, or
// the contiguous range stripTx0 .. stripTx0 + nStripTx - 1.
__device__ __forceinline__ void stripOf(const StencilLaunch& a, int g, int& tx, int& tyg)
{
    const int k = g % a.nStripTx;
    tyg = g / a.nStripTx;
    tx = a.edgeList ? a.edgeTx[k] : a.stripTx0 + k;
}

// Batched grids: one strip id covers all columns of a grid; both paths live in
// the same kernel.
template <typename T, int H, int V, int RPW, Xform X>
__device__ __forceinline__ void stencilStrip(const StencilLaunch& a, const T (&w)[H * V],
                                             T* __restrict__ tile, int g)
{
    const int tilesX = (a.nx + kTW - 1) / kTW;
    const int tx = g % tilesX, tyg = g / tilesX;
    if (stripIsInterior(a, tx)) stencilStripImpl<T, H, V, RPW, X, true>(a, w, tile, tx, tyg);
    else stencilStripImpl<T, H, V, RPW, X, false>(a, w, tile, tx, tyg);
}

}  // namespace dev
}  // namespace custen
