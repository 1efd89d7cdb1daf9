// kernels.h -- launch descriptors shared by the host plan layer and the gfx950
// stencil kernels.  Internal to the library.
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace custen {

enum class Dtype : int { F64 = 0, F32 = 1 };

// Element transform applied while staging input into LDS.
enum class Xform : int { Identity = 0, CubeMinus = 1 };

// Point functions the library ships (custenGetBuiltinFunction).
enum class Builtin : int {
    None = 0,
    XCentralDiff = 1,    // "x_central_difference"
    YCentralDiff9 = 2,   // "y_central_difference9"
    XYWeightedSum = 3,   // "xy_weighted_sum"
    XYCubeMinus = 4,     // "xy_cube_minus"
};

// Row addressing of the input region.
//   wrap:   source row = (gy mod rowWrap), used for periodic y on resident data.
//   linear: source row = clamp(gy, rowMin, rowMax), used for non-periodic y and
//           for halo-extended staging buffers (rowMin may be negative).
struct RowMap {
    int wrap;        // 1 = wrap, 0 = linear/clamp
    int rowWrap;     // period for wrap
    int rowMin;      // clamp bounds for linear
    int rowMax;
};

// One stencil apply over a rectangular region of rows [0, ny) of a grid that
// is nx wide.  Output row gy is written at out + gy * ldOut.
struct StencilLaunch {
    Dtype dtype;
    const void* in;          // element (0,0) of the region
    void* out;
    const double* weights;   // H*V weights (device-accessible), row-major
    int64_t ldIn, ldOut;     // row strides in elements
    int nx, ny;              // region size
    int H, L, V, T;          // stencil extents: R = H-1-L, B = V-1-T
    int xWrap;               // 1 periodic in x, 0 clamp
    RowMap rows;
    // write window in region coordinates (x) and global coordinates (y)
    int wx0, wx1;            // columns [wx0, wx1) written with the sum
    int wy0, wy1;            // global rows [wy0, wy1) written
    int yGlobal0;            // global row index of region row 0
    int zeroRight;           // columns [nx-R, nx) written as 0 (x-np)
    Xform xform;
    int rowsPerWave;         // 0 = default
    int chunks;              // row chunks per workgroup, 0 = default
    int colsPerLane;         // 1 or 2 (0 = default 2)
    int vecIn, vecOut;       // set by launchStencil: rows are pair-aligned
    // column strips handled by one launch: tx = stripTx0 .. stripTx0 +
    // nStripTx - 1, or (edgeList = 1) the nStripTx listed in edgeTx
    int stripTx0, nStripTx;
    int edgeList;
    int edgeTx[4];
    // 1: interior strips run the band walker (identity transform): one wave
    // per rowsPerWave * chunks rows of a strip
    int band;
};

// Point-function apply.  Same geometry as StencilLaunch; `func` is a device
// function pointer (indirect path) or null with `builtin` set (inlined path).
struct FunLaunch {
    const double* in;
    double* out;
    const double* coe;       // numCoe coefficients (device-accessible)
    int numCoe;
    const void* func;        // device function pointer
    Builtin builtin;         // inlined functor if != None
    int dir;                 // 0 x, 1 y, 2 xy
    int64_t ldIn, ldOut;
    int nx, ny;
    int H, L, V, T;
    int xWrap;
    RowMap rows;
    int wx0, wx1, wy0, wy1, yGlobal0;
    int rowsPerWave, chunks, colsPerLane;
    int vecIn;
    int band;                // as StencilLaunch::band (weighted-sum built-ins)
};

hipError_t launchStencil(const StencilLaunch& s, hipStream_t stream);
hipError_t launchFun(const FunLaunch& f, hipStream_t stream);

// Device addresses of the built-in point functions on the current device.
hipError_t builtinFunctionAddress(Builtin b, void** addr);

// LDS bytes the stencil kernel uses for a given shape (reported in mem_shared).
int stencilLdsBytes(Dtype dt, int H, int V, int rowsPerWave);
// LDS bytes of one band-walker workgroup (one wave).
int stencilBandLdsBytes(Dtype dt, int H, int V, int rowsPerWave);

// WENO5 upwind advection over a region of rows [0, ny), periodic in x: out =
// u * Fx + v * Fy.  `in` rows are addressed through `rows` (the kernel reads
// region rows -3 .. ny + 2); u, v and out are region-local.
struct WenoLaunch {
    const double* in;
    const double* u;
    const double* v;
    double* out;
    int64_t ldIn, ldSide, ldOut;  // row strides of in, of u and v, of out
    int nx, ny;
    RowMap rows;
    double coeDx, coeDy;          // 1/dx, 1/dy
    int singleSquares;            // 1: smoothness squares in single precision
    int rowsPerWave;              // 4 or 8 (0 = default)
};

hipError_t launchWeno(const WenoLaunch& w, hipStream_t stream);
int wenoLdsBytes(int rowsPerWave);

}  // namespace custen
