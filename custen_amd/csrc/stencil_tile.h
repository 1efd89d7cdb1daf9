// stencil_tile.h -- device building blocks of the gfx950 stencil kernels.
//
// Workgroup shape: 256 threads = 4 wave64s.  A workgroup owns a column strip
// kTW = 128 wide (each lane owns two adjacent columns, so one wave covers a
// row with one 16-byte load per lane for FP64) and walks down it in chunks of
// TH = 4 * RPW rows.  For each chunk the input rows [y0-T, y0+TH+B) and
// columns [x0-L, x0+kTW+R) are staged once into LDS; periodic wrap is resolved
// while forming the source address of each staged element, so the compute
// loop never branches on the boundary.  Each wave then produces RPW output
// rows x 2 columns per lane with a register sliding window over the V input
// rows that touch them.
//
// Summation order per output point: fma chain from 0.0, stencil row j outer,
// stencil column i inner (the order of cuSten/src/kernels/2d_xy_p_kernel.cu:505-520).
// This file must be compiled with -ffp-contract=off; the only fused operations
// are the explicit fma() calls.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

constexpr int kLanes = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kLanes * kWaves;
constexpr int kTW = 2 * kLanes;  // columns per workgroup strip

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = f64x2; };
template <> struct Vec2<float> { using type = f32x2; };

__device__ __forceinline__ int wrapIndex(int i, int n)
{
    if (i >= 0 && i < n) return i;
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int mapRow(const RowMap& m, int gy)
{
    if (m.wrap) return wrapIndex(gy, m.rowWrap);
    return gy < m.rowMin ? m.rowMin : (gy > m.rowMax ? m.rowMax : gy);
}

__device__ __forceinline__ int mapCol(int gx, int nx, int xWrap)
{
    if (xWrap) return wrapIndex(gx, nx);
    return gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx);
}

template <typename T, Xform X>
__device__ __forceinline__ T xformValue(T u)
{
    if constexpr (X == Xform::CubeMinus) {
        // u^3 - u evaluated as ((u*u)*u) - u without fusion.
        T u2 = u * u;
        T u3 = u2 * u;
        return u3 - u;
    } else {
        return u;
    }
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Workgroup id -> strip id, so that consecutive strips land on the same XCD
// (the hardware deals workgroups round-robin over the 8 XCDs, each with its
// own L2).
__device__ __forceinline__ int xcdRemap(int g, int G)
{
    const int per = G >> 3;
    const int full = per << 3;
    return g < full ? (g & 7) * per + (g >> 3) : g;
}

// LDS row geometry for a stencil of width H with left extent L:
// LDS column k holds global column x0 - LP + k, where LP = L rounded up to
// even so that the strip body starts 16-byte aligned.  A strip is 64 lanes x
// CPL columns wide (CPL = columns per lane, 1 or 2).
__host__ __device__ constexpr int ldsRowLengthW(int TW, int H) { return (TW + H + 1) & ~1; }
__host__ __device__ constexpr int ldsRowLength(int H) { return ldsRowLengthW(kTW, H); }

// The per-lane element group of a strip row: one value (CPL = 1) or an
// aligned pair (CPL = 2).
// Global-memory (address space 1) views.  Loads and stores through them are
// global rather than flat instructions even when the pointer was read from a
// table in memory (batched grids); flat operations would also count against
// the LDS wait counter, so every LDS wait would drain the row prefetch.
template <typename T> using Glob = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ const Glob<T>* globalPtr(const void* p)
{
    return (const Glob<T>*)p;
}
template <typename T>
__device__ __forceinline__ Glob<T>* globalPtr(void* p)
{
    return (Glob<T>*)p;
}

template <typename T, int CPL> struct LaneVec;
template <typename T> struct LaneVec<T, 2> {
    using type = typename Vec2<T>::type;
    __device__ static T get(const type& v, int e) { return e == 0 ? v.x : v.y; }
    __device__ static void set(type& v, int e, T x)
    {
        if (e == 0) v.x = x;
        else v.y = x;
    }
};
template <typename T> struct LaneVec<T, 1> {
    using type = T;
    __device__ static T get(const type& v, int) { return v; }
    __device__ static void set(type& v, int, T x) { v = x; }
};

// Stage one input row (all of [x0-L, x0+kTW+R)) into LDS.  One wave per row.
template <typename T, Xform X>
__device__ __forceinline__ void stageRow(T* __restrict__ lrow, const T* __restrict__ grow, int x0,
                                         int nx, int L, int R, int LP, int xWrap, bool vecOK,
                                         int lane)
{
    using V2 = typename Vec2<T>::type;
    const int gx = x0 + 2 * lane;
    if (vecOK) {
        V2 v = *reinterpret_cast<const V2*>(grow + gx);
        V2 o;
        o.x = xformValue<T, X>(v.x);
        o.y = xformValue<T, X>(v.y);
        *reinterpret_cast<V2*>(lrow + LP + 2 * lane) = o;
    } else {
        lrow[LP + 2 * lane] = xformValue<T, X>(grow[mapCol(gx, nx, xWrap)]);
        lrow[LP + 2 * lane + 1] = xformValue<T, X>(grow[mapCol(gx + 1, nx, xWrap)]);
    }
    if (lane < L) {
        lrow[LP - L + lane] = xformValue<T, X>(grow[mapCol(x0 - L + lane, nx, xWrap)]);
    } else if (lane < L + R) {
        const int k = lane - L;
        lrow[LP + kTW + k] = xformValue<T, X>(grow[mapCol(x0 + kTW + k, nx, xWrap)]);
    }
}

// Stage rows [y0 - T, y0 - T + nrows) of the region into LDS rows [0, nrows).
template <typename T, Xform X>
__device__ __forceinline__ void stageTile(T* __restrict__ tile, int LD, int nrows,
                                          const T* __restrict__ in, int64_t ldIn,
                                          const RowMap& rows, int y0, int T_, int x0, int nx,
                                          int L, int R, int LP, int xWrap, bool vecIn, int wave,
                                          int lane)
{
    for (int r = wave; r < nrows; r += kWaves) {
        const int sy = mapRow(rows, y0 - T_ + r);
        stageRow<T, X>(tile + r * LD, in + (int64_t)sy * ldIn, x0, nx, L, R, LP, xWrap, vecIn,
                       lane);
    }
}

// Rows of one chunk held in registers between their global load and their
// LDS store, so that a wave issues all its row loads before it waits on any
// of them (and, with prefetch, while the previous chunk is being computed).
template <typename T, int NR, int CPL = 2>
struct StagedRows {
    static constexpr int PER = (NR + kWaves - 1) / kWaves;  // rows per wave
    typename LaneVec<T, CPL>::type body[PER];
    // halo columns of the wave's PER rows, one value per lane: lane
    // i * (L + R) + k holds column k of the wave's i-th row (PER * (L + R)
    // values; a second register covers up to 128 of them)
    T halo[2];
};

// Source row of region row gy when the region is at least as tall as the
// window (one correction at most); the general mapRow otherwise.
__device__ __forceinline__ int mapRowNear(const RowMap& m, int gy)
{
    if (m.wrap) {
        if (gy < 0) gy += m.rowWrap;
        else if (gy >= m.rowWrap) gy -= m.rowWrap;
        return gy;
    }
    return gy < m.rowMin ? m.rowMin : (gy > m.rowMax ? m.rowMax : gy);
}

// FAST: the strip is a full 128 columns of pair-aligned rows, and row and
// halo-column indices need at most one wrap (ny >= window height, nx >= 128
// > any extent): no division in the address arithmetic.
template <typename T, int NR, bool FAST, int CPL = 2>
__device__ __forceinline__ void loadRows(StagedRows<T, NR, CPL>& s, const Glob<T>* __restrict__ in,
                                         int64_t ldIn, const RowMap& rows, int y0, int T_, int x0,
                                         int nx, int L, int R, int xWrap, bool vecIn, int wave,
                                         int lane)
{
    using LV = LaneVec<T, CPL>;
    using VT = typename LV::type;
    constexpr int TW = kLanes * CPL;
#pragma unroll
    for (int i = 0; i < StagedRows<T, NR, CPL>::PER; i++) {
        const int r = wave + i * kWaves;
        if (r < NR) {
            const int sy = FAST ? mapRowNear(rows, y0 - T_ + r) : mapRow(rows, y0 - T_ + r);
            const Glob<T>* grow = in + (int64_t)sy * ldIn;
            const int gx = x0 + CPL * lane;
            if (FAST || vecIn) {
                s.body[i] = *reinterpret_cast<const Glob<VT>*>(grow + gx);
            } else {
#pragma unroll
                for (int e = 0; e < CPL; e++) LV::set(s.body[i], e, grow[mapCol(gx + e, nx, xWrap)]);
            }
        }
    }
    // halo columns: lane q -> (row i, column k) with q = i * (L + R) + k
    const int HW = L + R;
    if (HW > 0) {
#pragma unroll
        for (int part = 0; part < 2; part++) {
            const int q = lane + part * kLanes;
            const int i = q / HW;
            const int k = q - i * HW;
            const int r = wave + i * kWaves;
            if (i < StagedRows<T, NR, CPL>::PER && r < NR) {
                const int sy = FAST ? mapRowNear(rows, y0 - T_ + r) : mapRow(rows, y0 - T_ + r);
                int hx = k < L ? x0 - L + k : x0 + TW + (k - L);
                if (FAST) {
                    if (xWrap) hx = hx < 0 ? hx + nx : (hx >= nx ? hx - nx : hx);
                    else hx = hx < 0 ? 0 : (hx >= nx ? nx - 1 : hx);
                } else {
                    hx = mapCol(hx, nx, xWrap);
                }
                s.halo[part] = in[(int64_t)sy * ldIn + hx];
            }
        }
    }
}

template <typename T, Xform X, int NR, int CPL = 2>
__device__ __forceinline__ void storeRows(const StagedRows<T, NR, CPL>& s, T* __restrict__ tile,
                                          int LD, int L, int R, int LP, int wave, int lane)
{
    using LV = LaneVec<T, CPL>;
    using VT = typename LV::type;
    constexpr int TW = kLanes * CPL;
#pragma unroll
    for (int i = 0; i < StagedRows<T, NR, CPL>::PER; i++) {
        const int r = wave + i * kWaves;
        if (r < NR) {
            T* lrow = tile + r * LD;
            VT o;
#pragma unroll
            for (int e = 0; e < CPL; e++) LV::set(o, e, xformValue<T, X>(LV::get(s.body[i], e)));
            *reinterpret_cast<VT*>(lrow + LP + CPL * lane) = o;
        }
    }
    const int HW = L + R;
    if (HW > 0) {
#pragma unroll
        for (int part = 0; part < 2; part++) {
            const int q = lane + part * kLanes;
            const int i = q / HW;
            const int k = q - i * HW;
            const int r = wave + i * kWaves;
            if (i < StagedRows<T, NR, CPL>::PER && r < NR) {
                const int hx = k < L ? LP - L + k : LP + TW + (k - L);
                tile[r * LD + hx] = xformValue<T, X>(s.halo[part]);
            }
        }
    }
}

// Write one output row's CPL values for this lane, honouring the write window.
template <typename T, int CPL>
__device__ __forceinline__ void storeLane(Glob<T>* __restrict__ orow, int gx, const T* v,
                                          bool fullRow, int nx, int wx0, int wx1, int zeroRight,
                                          int R)
{
    using LV = LaneVec<T, CPL>;
    using VT = typename LV::type;
    if (fullRow) {
        VT o;
#pragma unroll
        for (int e = 0; e < CPL; e++) LV::set(o, e, v[e]);
        __builtin_nontemporal_store(o, reinterpret_cast<Glob<VT>*>(orow + gx));
        return;
    }
#pragma unroll
    for (int e = 0; e < CPL; e++) {
        const int x = gx + e;
        if (x >= nx) continue;
        if (x >= wx0 && x < wx1) orow[x] = v[e];
        else if (zeroRight && x >= nx - R) orow[x] = T(0);
    }
}

// Write one output row's two values for this lane, honouring the write window.
template <typename T>
__device__ __forceinline__ void storePair(T* __restrict__ orow, int gx, T v0, T v1, bool fullRow,
                                          int nx, int wx0, int wx1, int zeroRight, int R)
{
    using V2 = typename Vec2<T>::type;
    if (fullRow) {
        V2 v;
        v.x = v0;
        v.y = v1;
        __builtin_nontemporal_store(v, reinterpret_cast<V2*>(orow + gx));
        return;
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int x = gx + e;
        const T val = e == 0 ? v0 : v1;
        if (x >= nx) continue;
        if (x >= wx0 && x < wx1) orow[x] = val;
        else if (zeroRight && x >= nx - R) orow[x] = T(0);
    }
}

}  // namespace dev
}  // namespace custen
