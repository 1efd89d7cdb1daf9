// stencil_kernels.hip -- gfx950 stencil kernels behind cuStenCompute2D*.
//
//   stencilFixed   : compile-time stencil shape (H x V), register sliding
//                    window, weights held in registers.  Serves the common
//                    x / y / xy shapes and the built-in xy point functions.
//   stencilGeneric : any shape up to 65 x 65, weights staged in LDS.
//   funKernel      : point-function ("Fun") variants; calls the user's device
//                    function through its pointer, or an inlined built-in.
//
// Semantics follow the reference kernels (cuSten/src/kernels/2d_*_kernel.cu):
// out(y,x) = fma chain over j < V, i < H of w[j*H+i] * in(y-T+j, x-L+i),
// starting from 0.0; periodic variants wrap, non-periodic variants write only
// the points whose whole window lies in the domain (see StencilLaunch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "kernels.h"
#include "stencil_band.h"
#include "stencil_body.h"
#include "stencil_tile.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

// ---------------------------------------------------------------------------
// Fixed-shape kernel
// ---------------------------------------------------------------------------
// Occupancy target per shape: the register-blocked window of the larger
// stencils is kept within 3 waves per SIMD without spilling.
#ifndef CUSTEN_OCC_F64_LARGE
#define CUSTEN_OCC_F64_LARGE 3
#endif
template <typename T, int H, int V, int RPW, int CPL>
struct FixedOccupancy {
    static constexpr int value =
        (sizeof(T) == 8 && H * V >= 49) ? 2
        : (sizeof(T) == 8 && H * V >= 25 && RPW == 4 && CPL == 2) ? CUSTEN_OCC_F64_LARGE
                                                                  : 1;
};

template <typename T, int H, int V, int RPW, Xform X, bool FAST, int CPL>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(FixedOccupancy<T, H, V, RPW, CPL>::value, 8))) void stencilFixed(
    StencilLaunch a)
{
    __shared__ __attribute__((aligned(16)))
    T tile[(kWaves * RPW + V - 1) * ldsRowLengthW(kLanes * CPL, H)];
    T w[H * V];
#pragma unroll
    for (int k = 0; k < H * V; k++) w[k] = static_cast<T>(a.weights[k]);
    int tx, tyg;
    stripOf(a, xcdRemap(blockIdx.x, gridDim.x), tx, tyg);
    stencilStripImpl<T, H, V, RPW, X, FAST, CPL>(a, w, tile, tx, tyg);
}

// ---------------------------------------------------------------------------
// Generic-shape kernel (runtime H, V); LDS = tile + weights.
// ---------------------------------------------------------------------------
constexpr int kGenericRPW = 4;

template <typename T, Xform X>
__global__ __launch_bounds__(kThreads) void stencilGeneric(StencilLaunch a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int RPW = kGenericRPW;
    constexpr int TH = kWaves * RPW;
    const int H = a.H, V = a.V;
    const int NR = TH + V - 1;
    const int LD = ldsRowLength(H);
    T* tile = reinterpret_cast<T*>(smem);
    T* wl = tile + NR * LD;

    const T* __restrict__ in = static_cast<const T*>(a.in);
    T* __restrict__ out = static_cast<T*>(a.out);

    for (int k = threadIdx.x; k < H * V; k += kThreads) wl[k] = static_cast<T>(a.weights[k]);

    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int L = a.L;
    const int R = H - 1 - L;
    const int LP = (L + 1) & ~1;
    const int nx = a.nx;
    const int tilesX = (nx + kTW - 1) / kTW;
    const int g = xcdRemap(blockIdx.x, gridDim.x);
    const int tx = g % tilesX;
    const int tyg = g / tilesX;
    const int x0 = tx * kTW;
    const bool inside = x0 + kTW <= nx;
    const bool vecIn = a.vecIn && inside;
    const bool fullRow = a.vecOut && inside && x0 >= a.wx0 && x0 + kTW <= a.wx1;

    for (int c = 0; c < a.chunks; c++) {
        const int y0 = (tyg * a.chunks + c) * TH;
        if (y0 >= a.ny) break;
        __syncthreads();
        stageTile<T, X>(tile, LD, NR, in, a.ldIn, a.rows, y0, a.T, x0, nx, L, R, LP, a.xWrap,
                        vecIn, wave, lane);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPW; r++) {
            const int gy = y0 + wave * RPW + r;
            if (gy >= a.ny) break;
            const int gyg = a.yGlobal0 + gy;
            if (gyg < a.wy0 || gyg >= a.wy1) continue;
            T s0 = T(0), s1 = T(0);
            const T* base = tile + (wave * RPW + r) * LD + (LP - L) + 2 * lane;
            for (int j = 0; j < V; j++) {
                const T* rowp = base + j * LD;
                for (int i = 0; i < H; i++) {
                    const T wk = wl[j * H + i];
                    s0 = fmaT(wk, rowp[i], s0);
                    s1 = fmaT(wk, rowp[i + 1], s1);
                }
            }
            storePair<T>(out + (int64_t)gy * a.ldOut, x0 + 2 * lane, s0, s1, fullRow, nx, a.wx0,
                         a.wx1, a.zeroRight, R);
        }
    }
}

// ---------------------------------------------------------------------------
// Point-function kernels
// ---------------------------------------------------------------------------
typedef double (*FnX)(double*, double*, int);
typedef double (*FnY)(double*, double*, int, int);
typedef double (*FnXY)(double*, double*, int, int, int, int);

constexpr int kFunRPW = 4;
constexpr int kMaxCoe = 1024;

struct CallIndirect {
    template <int DIR>
    __device__ static double call(const void* fn, double* d, double* c, int loc, int jump, int H,
                                  int V)
    {
        if constexpr (DIR == 0) return reinterpret_cast<FnX>(const_cast<void*>(fn))(d, c, loc);
        else if constexpr (DIR == 1)
            return reinterpret_cast<FnY>(const_cast<void*>(fn))(d, c, loc, jump);
        else return reinterpret_cast<FnXY>(const_cast<void*>(fn))(d, c, loc, jump, H, V);
    }
};

// Built-in point functions.  The inlined functors and the addressable device
// functions below share these bodies, so both paths give identical bits.
__device__ __forceinline__ double bodyXCentralDiff(const double* d, const double* c, int l)
{
    double t = d[l - 1] - 2.0 * d[l];
    t = t + d[l + 1];
    return t * c[0];
}
__device__ __forceinline__ double bodyYCentralDiff9(const double* d, const double* c, int l, int jump)
{
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) r = __builtin_fma(c[i], d[(l - 4 * jump) + i * jump], r);
    return r;
}
__device__ __forceinline__ double bodyXYWeightedSum(const double* d, const double* c, int l,
                                                    int jump, int nx, int ny)
{
    double r = 0.0;
    int cnt = 0;
    for (int j = 0; j < ny; j++) {
        const int t = l + j * jump;
        for (int i = 0; i < nx; i++) r = __builtin_fma(c[cnt++], d[t + i], r);
    }
    return r;
}
__device__ __forceinline__ double bodyXYCubeMinus(const double* d, const double* c, int l, int jump,
                                                  int nx, int ny)
{
    double r = 0.0;
    int cnt = 0;
    for (int j = 0; j < ny; j++) {
        const int t = l + j * jump;
        for (int i = 0; i < nx; i++) {
            const double u = d[t + i];
            const double u2 = u * u;
            const double u3 = u2 * u;
            r = __builtin_fma(c[cnt++], u3 - u, r);
        }
    }
    return r;
}

struct CallXCentralDiff {
    template <int DIR>
    __device__ static double call(const void*, double* d, double* c, int loc, int, int, int)
    {
        return bodyXCentralDiff(d, c, loc);
    }
};
struct CallYCentralDiff9 {
    template <int DIR>
    __device__ static double call(const void*, double* d, double* c, int loc, int jump, int, int)
    {
        return bodyYCentralDiff9(d, c, loc, jump);
    }
};

template <int DIR, class F>
__global__ __launch_bounds__(kThreads) void funKernel(FunLaunch a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int RPW = kFunRPW;
    constexpr int TH = kWaves * RPW;
    const int H = a.H, V = a.V;
    const int NR = TH + V - 1;
    const int LD = ldsRowLength(H);
    double* tile = reinterpret_cast<double*>(smem);
    double* coeL = tile + NR * LD;

    for (int k = threadIdx.x; k < a.numCoe; k += kThreads) coeL[k] = a.coe[k];

    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int L = a.L;
    const int R = H - 1 - L;
    const int LP = (L + 1) & ~1;
    const int nx = a.nx;
    const int tilesX = (nx + kTW - 1) / kTW;
    const int g = xcdRemap(blockIdx.x, gridDim.x);
    const int tx = g % tilesX;
    const int tyg = g / tilesX;
    const int x0 = tx * kTW;
    const bool vecIn = a.vecIn && (x0 + kTW <= nx);

    for (int c = 0; c < a.chunks; c++) {
        const int y0 = (tyg * a.chunks + c) * TH;
        if (y0 >= a.ny) break;
        __syncthreads();
        stageTile<double, Xform::Identity>(tile, LD, NR, a.in, a.ldIn, a.rows, y0, a.T, x0, nx, L,
                                           R, LP, a.xWrap, vecIn, wave, lane);
        __syncthreads();
        for (int r = 0; r < RPW; r++) {
            const int gy = y0 + wave * RPW + r;
            if (gy >= a.ny) break;
            const int gyg = a.yGlobal0 + gy;
            if (gyg < a.wy0 || gyg >= a.wy1) continue;
            const int rr = wave * RPW + r;
            for (int e = 0; e < 2; e++) {
                const int gx = x0 + 2 * lane + e;
                if (gx >= nx || gx < a.wx0 || gx >= a.wx1) continue;
                // window top-left in LDS
                const int tl = rr * LD + (LP - L) + 2 * lane + e;
                int loc;
                if constexpr (DIR == 0) loc = tl + L;           // centre (V == 1)
                else if constexpr (DIR == 1) loc = tl + a.T * LD;  // centre (H == 1)
                else loc = tl;                                  // top-left
                const double val = F::template call<DIR>(a.func, tile, coeL, loc, LD, H, V);
                a.out[(int64_t)gy * a.ldOut + gx] = val;
            }
        }
    }
}

// Point functions reached through a pointer (the caller's own device
// functions).  The call cannot be inlined, and the callee reads its window
// and coefficients through generic pointers: every read is a FLAT
// instruction, and the texture-address unit that processes them is the
// bottleneck (measured: busy 96 % of the kernel's time, about 21 cycles per
// 64-lane 8-byte read; a 3 x 3 weighted sum issues 18 of them per point).
// The kernel therefore spends as few of those units' cycles as it can on
// its own loads:
//   - one wave per workgroup walks a 128-column strip (two columns per
//     lane) over a band of kCallRows rows; each staged row is one 16-byte
//     load per lane plus one load for the halo columns, so staging costs
//     about one address instruction per 64 points;
//   - every lane calls for every row of the band (lanes past the grid or the
//     write window compute on staged values and do not store), so the call
//     sits in straight-line code with a full execution mask;
//   - results stay in registers until the band's last call: a callee starts
//     by waiting for all of the wave's outstanding memory operations, stores
//     included.
constexpr int kCallRows = 8;

template <int DIR>
__global__ __launch_bounds__(kLanes) void funCallKernel(FunLaunch a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int H = a.H, V = a.V, L = a.L;
    const int R = H - 1 - L;
    const int LP = (L + 1) & ~1;
    const int LD = ldsRowLength(H);
    const int NR = kCallRows + V - 1;
    double* tile = reinterpret_cast<double*>(smem);
    double* coeL = tile + NR * LD;
    const int lane = threadIdx.x;
    const int nx = a.nx;
    const int stripsX = (nx + kTW - 1) / kTW;
    const int g = xcdRemap(blockIdx.x, gridDim.x);
    const int x0 = (g % stripsX) * kTW;
    const int y0 = (g / stripsX) * kCallRows;
    if (y0 >= a.ny) return;

    for (int k = lane; k < a.numCoe; k += kLanes) coeL[k] = a.coe[k];
    const bool vecIn = a.vecIn && (x0 + kTW <= nx);
    for (int r = 0; r < NR; r++) {
        const int sy = mapRow(a.rows, y0 - a.T + r);
        stageRow<double, Xform::Identity>(tile + r * LD, a.in + (int64_t)sy * a.ldIn, x0, nx, L,
                                          R, LP, a.xWrap, vecIn, lane);
    }
    __syncthreads();

    // window of column x0 + 2 lane + e, row r: top-left at r LD + base + e
    int base = (LP - L) + 2 * lane;
    if constexpr (DIR == 0) base += L;           // window centre (V == 1)
    else if constexpr (DIR == 1) base += a.T * LD;  // window centre (H == 1)
    const int rows = min(kCallRows, a.ny - y0);
    const int gyg0 = a.yGlobal0 + y0;
    double res[kCallRows][2];
#pragma unroll
    for (int r = 0; r < kCallRows; r++) {
        res[r][0] = res[r][1] = 0.0;
        if (r < rows && gyg0 + r >= a.wy0 && gyg0 + r < a.wy1) {
            res[r][0] = CallIndirect::call<DIR>(a.func, tile, coeL, base + r * LD, LD, H, V);
            res[r][1] = CallIndirect::call<DIR>(a.func, tile, coeL, base + r * LD + 1, LD, H, V);
        }
    }
    const int gx = x0 + 2 * lane;
    double* orow = a.out + (int64_t)y0 * a.ldOut + gx;
#pragma unroll
    for (int r = 0; r < kCallRows; r++) {
        if (r < rows && gyg0 + r >= a.wy0 && gyg0 + r < a.wy1) {
#pragma unroll
            for (int e = 0; e < 2; e++)
                if (gx + e < nx && gx + e >= a.wx0 && gx + e < a.wx1)
                    orow[(int64_t)r * a.ldOut + e] = res[r][e];
        }
    }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Addressable built-in functions (custenGetBuiltinFunction)
// ---------------------------------------------------------------------------
__device__ double custenBuiltinXCentralDifference(double* d, double* c, int loc)
{
    return dev::bodyXCentralDiff(d, c, loc);
}
__device__ double custenBuiltinYCentralDifference9(double* d, double* c, int loc, int jump)
{
    return dev::bodyYCentralDiff9(d, c, loc, jump);
}
__device__ double custenBuiltinXYWeightedSum(double* d, double* c, int loc, int jump, int nx, int ny)
{
    return dev::bodyXYWeightedSum(d, c, loc, jump, nx, ny);
}
__device__ double custenBuiltinXYCubeMinus(double* d, double* c, int loc, int jump, int nx, int ny)
{
    return dev::bodyXYCubeMinus(d, c, loc, jump, nx, ny);
}

__device__ dev::FnX custenPtrXCentralDifference = custenBuiltinXCentralDifference;
__device__ dev::FnY custenPtrYCentralDifference9 = custenBuiltinYCentralDifference9;
__device__ dev::FnXY custenPtrXYWeightedSum = custenBuiltinXYWeightedSum;
__device__ dev::FnXY custenPtrXYCubeMinus = custenBuiltinXYCubeMinus;

hipError_t builtinFunctionAddress(Builtin b, void** addr)
{
    *addr = nullptr;
    switch (b) {
    case Builtin::XCentralDiff:
        return hipMemcpyFromSymbol(addr, HIP_SYMBOL(custenPtrXCentralDifference), sizeof(void*));
    case Builtin::YCentralDiff9:
        return hipMemcpyFromSymbol(addr, HIP_SYMBOL(custenPtrYCentralDifference9), sizeof(void*));
    case Builtin::XYWeightedSum:
        return hipMemcpyFromSymbol(addr, HIP_SYMBOL(custenPtrXYWeightedSum), sizeof(void*));
    case Builtin::XYCubeMinus:
        return hipMemcpyFromSymbol(addr, HIP_SYMBOL(custenPtrXYCubeMinus), sizeof(void*));
    default:
        return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
namespace {

constexpr int kDefaultRPW = 8;
constexpr int kDefaultChunks = 4;

template <typename T>
bool pairAligned(const void* p, int64_t ld)
{
    return (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T)) == 0) && (ld % 2 == 0);
}

int gridSize(int nx, int ny, int TH, int chunks)
{
    const int tilesX = (nx + dev::kTW - 1) / dev::kTW;
    const int rowsPerGroup = TH * chunks;
    const int tilesY = (ny + rowsPerGroup - 1) / rowsPerGroup;
    return tilesX * tilesY;
}

// Interior column strips run the FAST kernel (or the band walker), the edge
// strips the general one: by list when there are at most four, otherwise (no
// interior strip at all: unaligned rows, or a periodic grid lower than the
// stencil) as the contiguous ranges left and right of the interior.
template <typename T, int H, int V, int RPW, Xform X, int CPL = 2>
hipError_t runFixed(const StencilLaunch& in, hipStream_t st)
{
    StencilLaunch s = in;
    const int tw = dev::kLanes * CPL;
    const int tilesX = (s.nx + tw - 1) / tw;
    const int rowsPerGroup = dev::kWaves * RPW * s.chunks;
    const int tilesY = (s.ny + rowsPerGroup - 1) / rowsPerGroup;
    int lo = 0, hi = tilesX;  // interior range [lo, hi)
    while (lo < hi && !dev::stripIsInterior(s, lo, CPL)) lo++;
    while (hi > lo && !dev::stripIsInterior(s, hi - 1, CPL)) hi--;
    for (int t = lo; t < hi; t++)
        if (!dev::stripIsInterior(s, t, CPL)) return hipErrorInvalidValue;  // not contiguous
    // edge strips first, one row chunk per workgroup: they are a few percent
    // of the work and must not leave a tail after the interior launch
    const int tilesYe = (s.ny + dev::kWaves * RPW - 1) / (dev::kWaves * RPW);
    auto launchEdges = [&](const StencilLaunch& e) -> hipError_t {
        const int64_t groups = (int64_t)e.nStripTx * tilesYe;
        if (groups > 0x7fffffff) return hipErrorInvalidValue;
        hipLaunchKernelGGL((dev::stencilFixed<T, H, V, RPW, X, false, CPL>),
                           dim3((unsigned)groups), dim3(dev::kThreads), 0, st, e);
        return hipGetLastError();
    };
    const int nEdge = tilesX - (hi - lo);
    if (nEdge > 0 && nEdge <= 4) {
        StencilLaunch e = s;
        e.chunks = 1;
        e.edgeList = 1;
        e.nStripTx = 0;
        for (int t = 0; t < tilesX; t++)
            if (t < lo || t >= hi) e.edgeTx[e.nStripTx++] = t;
        hipError_t err = launchEdges(e);
        if (err != hipSuccess) return err;
    } else if (nEdge > 4) {
        const int ranges[2][2] = {{0, lo}, {hi, tilesX}};
        for (const auto& r : ranges) {
            if (r[1] <= r[0]) continue;
            StencilLaunch e = s;
            e.chunks = 1;
            e.edgeList = 0;
            e.stripTx0 = r[0];
            e.nStripTx = r[1] - r[0];
            hipError_t err = launchEdges(e);
            if (err != hipSuccess) return err;
        }
    }
    if (hi > lo) {
        s.edgeList = 0;
        s.stripTx0 = lo;
        s.nStripTx = hi - lo;
        // band walker: every weight stencil shape, and the cube-minus point
        // function on its xy shapes 3x3 and 5x5
        constexpr bool bandShape =
            X == Xform::Identity ||
            (X == Xform::CubeMinus && H == V && (H == 3 || H == 5) && sizeof(T) == 8);
        if constexpr (bandShape && CPL == 2 && (RPW == 4 || RPW == 8)) {
            // band walker: every row it touches must need at most one wrap
            if (s.band && (!s.rows.wrap || (s.rows.rowWrap >= s.V + RPW && s.ny <= s.rows.rowWrap))) {
                const int bandRows = RPW * s.chunks;
                const int64_t waves = (int64_t)s.nStripTx * ((s.ny + bandRows - 1) / bandRows);
                if (waves > 0x7fffffff) return hipErrorInvalidValue;
                // every walked point written, whole chunks: no store masks
                const bool full = lo * dev::kTW >= s.wx0 && hi * dev::kTW <= s.wx1 &&
                                  s.yGlobal0 >= s.wy0 && s.yGlobal0 + s.ny <= s.wy1 &&
                                  s.ny % RPW == 0;
                if (full || !dev::kStreamWeights<T, H, V, true>) {
                    if (full)
                        hipLaunchKernelGGL((dev::stencilBand<T, H, V, RPW, true, X>),
                                           dim3((unsigned)waves), dim3(dev::kLanes), 0, st, s);
                    else
                        hipLaunchKernelGGL((dev::stencilBand<T, H, V, RPW, false, X>),
                                           dim3((unsigned)waves), dim3(dev::kLanes), 0, st, s);
                    return hipGetLastError();
                }
                // Stencils whose masked walk keeps resident (spilled) weights:
                // the unmasked walk serves the strips whose every column is
                // written, [xa, xb), over the whole chunks of the written rows
                // [ya, ya + yFull); the masked walk serves the ragged rows and
                // the strips at the column window's edges.  Rows outside the
                // row window are not written and are skipped.  A row offset
                // needs clamped rows (wrapped rows are addressed from region
                // row 0).
                int xa = lo, xb = hi;
                while (xa < xb && xa * dev::kTW < s.wx0) xa++;
                while (xb > xa && xb * dev::kTW > s.wx1) xb--;
                const int ya = std::max(0, s.wy0 - s.yGlobal0);
                const int yb = std::min(s.ny, s.wy1 - s.yGlobal0);
                int yFull = yb > ya ? (yb - ya) / RPW * RPW : 0;
                if (s.rows.wrap && (ya != 0 || yFull != s.ny)) yFull = 0;
                auto walk = [&](bool fullWalk, const StencilLaunch& strips, int y0,
                                int n) -> hipError_t {
                    if (strips.nStripTx <= 0 || n <= 0) return hipSuccess;
                    StencilLaunch t = strips;
                    t.ny = n;
                    if (y0 > 0) {
                        t.in = static_cast<const char*>(s.in) + (int64_t)y0 * s.ldIn * sizeof(T);
                        t.out = static_cast<char*>(s.out) + (int64_t)y0 * s.ldOut * sizeof(T);
                        t.yGlobal0 = s.yGlobal0 + y0;
                        t.rows.rowMin = s.rows.rowMin - y0;
                        t.rows.rowMax = s.rows.rowMax - y0;
                    }
                    const unsigned w =
                        (unsigned)((int64_t)t.nStripTx * ((n + bandRows - 1) / bandRows));
                    if (fullWalk)
                        hipLaunchKernelGGL((dev::stencilBand<T, H, V, RPW, true, X>), dim3(w),
                                           dim3(dev::kLanes), 0, st, t);
                    else
                        hipLaunchKernelGGL((dev::stencilBand<T, H, V, RPW, false, X>), dim3(w),
                                           dim3(dev::kLanes), 0, st, t);
                    return hipGetLastError();
                };
                auto range = [&](int tx0, int ntx) {
                    StencilLaunch t = s;
                    t.edgeList = 0;
                    t.stripTx0 = tx0;
                    t.nStripTx = ntx;
                    return t;
                };
                if (yFull == 0 || xb <= xa) return walk(false, range(lo, hi - lo), 0, s.ny);
                hipError_t err;
                if ((err = walk(true, range(xa, xb - xa), ya, yFull)) != hipSuccess) return err;
                if ((err = walk(false, range(xa, xb - xa), ya + yFull, yb - ya - yFull)) !=
                    hipSuccess)
                    return err;
                const int nSide = (xa - lo) + (hi - xb);
                if (nSide <= 4) {
                    // both edges in one launch
                    StencilLaunch e = s;
                    e.edgeList = 1;
                    e.nStripTx = 0;
                    for (int tx = lo; tx < hi; tx++)
                        if (tx < xa || tx >= xb) e.edgeTx[e.nStripTx++] = tx;
                    return walk(false, e, ya, yb - ya);
                }
                if ((err = walk(false, range(lo, xa - lo), ya, yb - ya)) != hipSuccess) return err;
                return walk(false, range(xb, hi - xb), ya, yb - ya);
            }
        }
        hipLaunchKernelGGL((dev::stencilFixed<T, H, V, RPW, X, true, CPL>),
                           dim3(s.nStripTx * tilesY), dim3(dev::kThreads), 0, st, s);
        return hipGetLastError();
    }
    return hipSuccess;
}

template <typename T, Xform X>
hipError_t runGeneric(const StencilLaunch& s, hipStream_t st)
{
    const int TH = dev::kWaves * dev::kGenericRPW;
    const int grid = gridSize(s.nx, s.ny, TH, s.chunks);
    const size_t lds =
        sizeof(T) * ((size_t)(TH + s.V - 1) * dev::ldsRowLength(s.H) + (size_t)s.H * s.V);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dev::stencilGeneric<T, X>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((dev::stencilGeneric<T, X>), dim3(grid), dim3(dev::kThreads), lds, st, s);
    return hipGetLastError();
}

// Shapes with a compiled fixed kernel: x: H in {3,5,7,9}; y: V in {3,5,7,9};
// xy: 3x3, 5x5; and 1x1 (a scaled copy: the walkers' own memory ceiling).
// RPW in {4, 8}.
#define CUSTEN_FIXED_SHAPES(X_) \
    X_(3, 1) X_(5, 1) X_(7, 1) X_(9, 1) X_(1, 3) X_(1, 5) X_(1, 7) X_(1, 9) X_(3, 3) X_(5, 5) \
    X_(1, 1)
#define CUSTEN_WIDE_SHAPES(X_) \
    X_(11, 1) X_(13, 1) X_(1, 11) X_(1, 13) X_(7, 7) X_(9, 9) X_(3, 5) X_(5, 3)

template <typename T, Xform X>
hipError_t dispatchTyped(const StencilLaunch& s, hipStream_t st)
{
    const int rpw = s.rowsPerWave;
    // one column per lane (deeper row blocking) for the xy shapes
    if (s.colsPerLane == 1) {
        if (s.H == 5 && s.V == 5)
            return rpw == 16 ? runFixed<T, 5, 5, 16, X, 1>(s, st) : runFixed<T, 5, 5, 8, X, 1>(s, st);
        if (s.H == 3 && s.V == 3)
            return rpw == 16 ? runFixed<T, 3, 3, 16, X, 1>(s, st) : runFixed<T, 3, 3, 8, X, 1>(s, st);
    }
#define CUSTEN_CASE(HH, VV)                                                      \
    if (s.H == HH && s.V == VV) {                                                \
        if (rpw == 4) return runFixed<T, HH, VV, 4, X>(s, st);                   \
        return runFixed<T, HH, VV, 8, X>(s, st);                                 \
    }
    CUSTEN_FIXED_SHAPES(CUSTEN_CASE)
#undef CUSTEN_CASE
    // wider shapes (high-order schemes), identity transform, 4 rows per wave
    if constexpr (X == Xform::Identity) {
#define CUSTEN_WIDE(HH, VV) \
    if (s.H == HH && s.V == VV) return runFixed<T, HH, VV, 4, X>(s, st);
        CUSTEN_WIDE_SHAPES(CUSTEN_WIDE)
#undef CUSTEN_WIDE
    }
    return runGeneric<T, X>(s, st);
}

}  // namespace

int stencilLdsBytes(Dtype dt, int H, int V, int rowsPerWave)
{
    const int es = dt == Dtype::F64 ? 8 : 4;
    const int rpw = rowsPerWave > 0 ? rowsPerWave : kDefaultRPW;
    return es * (dev::kWaves * rpw + V - 1) * dev::ldsRowLength(H);
}

int stencilBandLdsBytes(Dtype dt, int H, int V, int rowsPerWave)
{
    const int rpw = rowsPerWave == 8 ? 8 : 4;
    const int d = 8 / rpw;
    const int hlp = H & ~1;  // H - 1 halo columns per side, rounded up to even
    return (dt == Dtype::F64 ? 8 : 4) * ((d + 1) * rpw + V - 1) * (dev::kTW + 2 * hlp);
}

hipError_t launchStencil(const StencilLaunch& in, hipStream_t st)
{
    StencilLaunch s = in;
    if (s.nx <= 0 || s.ny <= 0) return hipSuccess;
    if (s.H < 1 || s.V < 1 || s.H > 65 || s.V > 65) return hipErrorInvalidValue;
    if (s.colsPerLane != 1) s.colsPerLane = 2;
    if (s.colsPerLane == 1) {
        if (s.rowsPerWave != 8 && s.rowsPerWave != 16) s.rowsPerWave = 16;
    } else if (s.rowsPerWave != 4 && s.rowsPerWave != 8) {
        s.rowsPerWave = kDefaultRPW;
    }
    if (s.chunks <= 0) s.chunks = kDefaultChunks;
    if (s.dtype == Dtype::F64) {
        s.vecIn = pairAligned<double>(s.in, s.ldIn);
        s.vecOut = pairAligned<double>(s.out, s.ldOut);
        if (s.xform == Xform::CubeMinus) return dispatchTyped<double, Xform::CubeMinus>(s, st);
        return dispatchTyped<double, Xform::Identity>(s, st);
    }
    s.vecIn = pairAligned<float>(s.in, s.ldIn);
    s.vecOut = pairAligned<float>(s.out, s.ldOut);
    if (s.xform == Xform::CubeMinus) return hipErrorInvalidValue;
    return dispatchTyped<float, Xform::Identity>(s, st);
}

hipError_t launchFun(const FunLaunch& in, hipStream_t st)
{
    FunLaunch f = in;
    if (f.nx <= 0 || f.ny <= 0) return hipSuccess;
    if (f.H < 1 || f.V < 1 || f.H > 65 || f.V > 65) return hipErrorInvalidValue;
    if (f.numCoe < 0 || f.numCoe > dev::kMaxCoe) return hipErrorInvalidValue;
    if (f.chunks <= 0) f.chunks = kDefaultChunks;
    f.vecIn = pairAligned<double>(f.in, f.ldIn);

    // Built-in xy functions are weighted sums (of u or of u^3 - u): run them on
    // the register-blocked stencil kernel with the coefficients as weights.
    if (f.dir == 2 && (f.builtin == Builtin::XYWeightedSum || f.builtin == Builtin::XYCubeMinus) &&
        f.numCoe >= f.H * f.V) {
        StencilLaunch s{};
        s.dtype = Dtype::F64;
        s.in = f.in;
        s.out = f.out;
        s.weights = f.coe;
        s.ldIn = f.ldIn;
        s.ldOut = f.ldOut;
        s.nx = f.nx;
        s.ny = f.ny;
        s.H = f.H;
        s.L = f.L;
        s.V = f.V;
        s.T = f.T;
        s.xWrap = f.xWrap;
        s.rows = f.rows;
        s.wx0 = f.wx0;
        s.wx1 = f.wx1;
        s.wy0 = f.wy0;
        s.wy1 = f.wy1;
        s.yGlobal0 = f.yGlobal0;
        s.zeroRight = 0;
        s.xform = f.builtin == Builtin::XYCubeMinus ? Xform::CubeMinus : Xform::Identity;
        s.rowsPerWave = f.rowsPerWave;
        s.chunks = f.chunks;
        s.colsPerLane = f.colsPerLane;
        s.band = f.band;
        return launchStencil(s, st);
    }

    const int TH = dev::kWaves * dev::kFunRPW;
    const int grid = gridSize(f.nx, f.ny, TH, f.chunks);
    const size_t lds = sizeof(double) * ((size_t)(TH + f.V - 1) * dev::ldsRowLength(f.H) +
                                         (size_t)(f.numCoe > 0 ? f.numCoe : 1));
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        const void* kernels[] = {reinterpret_cast<const void*>(&dev::funKernel<0, dev::CallXCentralDiff>),
                                 reinterpret_cast<const void*>(&dev::funKernel<1, dev::CallYCentralDiff9>)};
        for (const void* k : kernels) {
            hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
    }

    if (f.builtin == Builtin::XCentralDiff && f.dir == 0) {
        hipLaunchKernelGGL((dev::funKernel<0, dev::CallXCentralDiff>), dim3(grid),
                           dim3(dev::kThreads), lds, st, f);
        return hipGetLastError();
    }
    if (f.builtin == Builtin::YCentralDiff9 && f.dir == 1) {
        hipLaunchKernelGGL((dev::funKernel<1, dev::CallYCentralDiff9>), dim3(grid),
                           dim3(dev::kThreads), lds, st, f);
        return hipGetLastError();
    }
    if (f.func == nullptr) return hipErrorInvalidValue;
    // caller's function through its pointer: one wave per 128 x kCallRows block
    const int64_t stripsX = (f.nx + dev::kTW - 1) / dev::kTW;
    const int64_t waves = stripsX * ((f.ny + dev::kCallRows - 1) / dev::kCallRows);
    if (waves > 0x7fffffff) return hipErrorInvalidValue;
    const size_t ldsCall =
        sizeof(double) * ((size_t)(dev::kCallRows + f.V - 1) * dev::ldsRowLength(f.H) +
                          (size_t)(f.numCoe > 0 ? f.numCoe : 1));
    if (ldsCall > 160 * 1024) return hipErrorInvalidValue;
    if (ldsCall > 64 * 1024) {
        const void* kernels[] = {reinterpret_cast<const void*>(&dev::funCallKernel<0>),
                                 reinterpret_cast<const void*>(&dev::funCallKernel<1>),
                                 reinterpret_cast<const void*>(&dev::funCallKernel<2>)};
        for (const void* k : kernels) {
            hipError_t e =
                hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsCall);
            if (e != hipSuccess) return e;
        }
    }
    switch (f.dir) {
    case 0:
        hipLaunchKernelGGL((dev::funCallKernel<0>), dim3((unsigned)waves), dim3(dev::kLanes),
                           ldsCall, st, f);
        break;
    case 1:
        hipLaunchKernelGGL((dev::funCallKernel<1>), dim3((unsigned)waves), dim3(dev::kLanes),
                           ldsCall, st, f);
        break;
    default:
        hipLaunchKernelGGL((dev::funCallKernel<2>), dim3((unsigned)waves), dim3(dev::kLanes),
                           ldsCall, st, f);
        break;
    }
    return hipGetLastError();
}

}  // namespace custen
