/*
 * cuSten.h -- public C ABI of the MI355X-native cuSten stencil library.
 *
 * Drop-in for the reference's umbrella header cuSten/cuSten.h, which exposes
 *   - the DEVICE / HOST offload macros              (cuSten/cuSten.h:30-31)
 *   - the handle type cuSten_t                       (cuSten/src/struct/cuSten_struct_type.h:84-122)
 *   - cuStenCreate/Swap/Destroy2D{X,Y,XY}{p,np}[Fun] (cuSten/src/struct/cuSten_struct_functions.h:67-838)
 *   - cuStenCompute2D{X,Y,XY}{p,np}[Fun]             (cuSten/src/kernels/stencil_kernels.h:52-178)
 *   - checkError                                     (cuSten/src/util/util.h:43)
 * with the same names, argument order and argument meaning.  Differences a
 * caller can observe are listed next to each declaration and in one place
 * here:
 *   - streams/events in cuSten_t are hipStream_t / hipEvent_t;
 *   - cuSten_t has one extra trailing field, `plan`, owned by the library;
 *   - all functions have C linkage (callable from C, C++ and ctypes);
 *   - Create validates its arguments; failures are recorded and readable with
 *     custenGetLastError() instead of being silently accepted;
 *   - data pointers may be device memory (hipMalloc), managed memory, or host
 *     memory (pinned or pageable).  Host data is streamed through HBM in
 *     numTiles strips; device data is computed in one pass.
 * Compute is asynchronous with respect to the host for device-resident data,
 * as in the reference (the caller synchronises, e.g. hipDeviceSynchronize());
 * for host-resident data it returns after the results are back in host memory.
 *
 * A handle is not re-entrant: use one handle per host thread.
 */
#ifndef CUSTEN_H
#define CUSTEN_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#define DEVICE 0
#define HOST 1

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque library state attached to a handle. */
typedef struct custenPlan custenPlan;

/* Handle, field for field the reference layout (cuSten_struct_type.h:84-122),
 * followed by `plan`.  The caller allocates it (usually on the stack) and the
 * library fills it in Create. */
typedef struct
{
    int deviceNum;
    int numStreams;
    int numTiles;
    int nx;
    int ny;
    int nyTile;
    int numSten;
    int numStenLeft;
    int numStenRight;
    int numStenTop;
    int numStenBottom;
    int numStenHoriz;
    int numStenVert;
    int BLOCK_X;
    int BLOCK_Y;
    int xGrid;
    int yGrid;
    int mem_shared;
    double** dataInput;
    double** dataOutput;
    double** uVel;
    double** vVel;
    double* weights;
    double* coe;
    double coeDx;
    double coeDy;
    int numCoe;
    int nxLocal;
    int nyLocal;
    double** boundaryTop;
    double** boundaryBottom;
    int numBoundaryTop;
    int numBoundaryBottom;
    hipStream_t* streams;
    hipEvent_t* events;
    double* devFunc;
    custenPlan* plan;
} cuSten_t;

/* User point-function signatures (2d_x_p_fun_kernel.cu:47, 2d_y_p_fun_kernel.cu:49,
 * 2d_xy_p_fun_kernel.cu:51).  data: the staged input window, coe: the
 * coefficients, loc: index of the window centre (x, y) or of the window's
 * top-left point (xy), jump: row stride of `data`, nx/ny: stencil width/height. */
typedef double (*devArg1X)(double*, double*, int);
typedef double (*devArg1Y)(double*, double*, int, int);
typedef double (*devArg1XY)(double*, double*, int, int, int, int);

/* ------------------------------------------------------------------------ */
/* x direction, weights.  cuSten_struct_functions.h:67-104 (np),            */
/* :184-221 (p); stencil_kernels.h:52 (np Compute), :63 (p Compute)          */
/* ------------------------------------------------------------------------ */
void cuStenCreate2DXnp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                       int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                       double* weights, int numSten, int numStenLeft, int numStenRight);
void cuStenSwap2DXnp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXnp(cuSten_t* pt_cuSten);
void cuStenCompute2DXnp(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DXp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                      int BLOCK_X, int BLOCK_Y, double* dataNew, double* dataOld,
                      double* weights, int numSten, int numStenLeft, int numStenRight);
void cuStenSwap2DXp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXp(cuSten_t* pt_cuSten);
void cuStenCompute2DXp(cuSten_t* pt_cuSten, bool offload);

/* x direction, point function.  cuSten_struct_functions.h:125-164 (np),
 * :242-289 (p); stencil_kernels.h:74 (np Compute), :85 (p Compute).  The reference defines the
 * x-periodic Compute only under the misspelt name cuSenCompute2DXpFun
 * (2d_x_p_fun_kernel.cu:172); both names are exported here. */
void cuStenCreate2DXnpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                          int BLOCK_X, int BLOCK_Y, double* dateOutput, double* dateInput,
                          double* coe, int numSten, int numStenLeft, int numStenRight,
                          int numCoe, double* func);
void cuStenSwap2DXnpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXnpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DXnpFun(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DXpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                         int BLOCK_X, int BLOCK_Y, double* dataNew, double* dataOld,
                         double* coe, int numSten, int numStenLeft, int numStenRight,
                         int numCoe, double* func);
void cuStenSwap2DXpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DXpFun(cuSten_t* pt_cuSten, bool offload);
void cuSenCompute2DXpFun(cuSten_t* pt_cuSten, bool offload);

/* ------------------------------------------------------------------------ */
/* y direction.  cuSten_struct_functions.h:626-663 (np), :742-779 (p),       */
/* :684-722 (np Fun), :801-840 (p Fun); stencil_kernels.h:96-133             */
/* ------------------------------------------------------------------------ */
void cuStenCreate2DYnp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                       int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                       double* weights, int numSten, int numStenTop, int numStenBottom);
void cuStenSwap2DYnp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DYnp(cuSten_t* pt_cuSten);
void cuStenCompute2DYnp(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DYp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                      int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                      double* weights, int numSten, int numStenTop, int numStenBottom);
void cuStenSwap2DYp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DYp(cuSten_t* pt_cuSten);
void cuStenCompute2DYp(cuSten_t* pt_cuSten, bool offload);

/* The reference's y-np Fun Create has no numCoe argument; it stages numSten
 * coefficients (2d_y_np_fun_kernel.cu:102). */
void cuStenCreate2DYnpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                          int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                          double* coe, int numSten, int numStenTop, int numStenBottom,
                          double* func);
void cuStenSwap2DYnpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DYnpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DYnpFun(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DYpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                         int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                         double* coe, int numSten, int numStenTop, int numStenBottom,
                         int numCoe, double* func);
void cuStenSwap2DYpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DYpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DYpFun(cuSten_t* pt_cuSten, bool offload);

/* ------------------------------------------------------------------------ */
/* xy (cross) stencils.  cuSten_struct_functions.h:371-411 (np),             */
/* :435-476 (np Fun), :500-540 (p), :564-605 (p Fun);                        */
/* stencil_kernels.h:140-173                                                 */
/* ------------------------------------------------------------------------ */
void cuStenCreate2DXYnp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                        int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                        double* weights, int numStenHoriz, int numStenLeft, int numStenRight,
                        int numStenVert, int numStenTop, int numStenBottom);
void cuStenSwap2DXYnp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXYnp(cuSten_t* pt_cuSten);
void cuStenCompute2DXYnp(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DXYnpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                           int BLOCK_X, int BLOCK_Y, double* dataNew, double* dataOld,
                           double* coe, int numStenHoriz, int numStenLeft, int numStenRight,
                           int numStenVert, int numStenTop, int numStenBottom, double* func);
void cuStenSwap2DXYnpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXYnpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DXYnpFun(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DXYp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                       int BLOCK_X, int BLOCK_Y, double* dataNew, double* dataOld,
                       double* weights, int numStenHoriz, int numStenLeft, int numStenRight,
                       int numStenVert, int numStenTop, int numStenBottom);
void cuStenSwap2DXYp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXYp(cuSten_t* pt_cuSten);
void cuStenCompute2DXYp(cuSten_t* pt_cuSten, bool offload);

void cuStenCreate2DXYpFun(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                          int BLOCK_X, int BLOCK_Y, double* dataOutput, double* dataInput,
                          double* coe, int numStenHoriz, int numStenLeft, int numStenRight,
                          int numStenVert, int numStenTop, int numStenBottom, double* func);
void cuStenSwap2DXYpFun(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXYpFun(cuSten_t* pt_cuSten);
void cuStenCompute2DXYpFun(cuSten_t* pt_cuSten, bool offload);

/* ---- 2D xy periodic WENO5 upwind advection ------------------------------- */
/* cuSten_struct_functions.h:309-346, stencil_kernels.h:184; behaviour
 * custenCreateDestroy2DXYADVWENOp.cu:57-260 and 2d_xyADVWENO_p_kernel.cu.
 * dataOutput = u * d(dataInput)/dx + v * d(dataInput)/dy with fifth-order WENO
 * fluxes upwinded on the signs of u and v; coeDx = 1/dx, coeDy = 1/dy.  u and
 * v are nx*ny fields like the data.  The smoothness squares are taken in
 * single precision by default, as the reference's powf calls on doubles do;
 * custenSetWenoSquares(h, 0) takes them in double. */
void cuStenCreate2DXYWENOADVp(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                              int BLOCK_X, int BLOCK_Y, double dx, double dy, double* u,
                              double* v, double* dataOutput, double* dataInput);
void cuStenSwap2DXYWENOADVp(cuSten_t* pt_cuSten, double* dataInput);
void cuStenDestroy2DXYWENOADVp(cuSten_t* pt_cuSten);
void cuStenCompute2DXYWENOADVp(cuSten_t* pt_cuSten, bool offload);

/* Error convention of the reference (util.h:43, error.cu:41-50): if the last
 * HIP call (or the last library call) failed, print `msg` with the error text
 * and exit(EXIT_FAILURE). */
void checkError(const char* msg);

/* ------------------------------------------------------------------------ */
/* Lower-case aliases (custenCreate2DXYp, custenCompute2DXYp, ...): every     */
/* cuSten* entry point above is also exported as custen* with the same        */
/* signature.  Declared here through one macro list.                          */
/* ------------------------------------------------------------------------ */

/* ======================================================================== */
/* Extensions (no counterpart in the reference)                              */
/* ======================================================================== */

/* Error codes recorded by every entry point. */
enum {
    CUSTEN_SUCCESS = 0,
    CUSTEN_ERROR_INVALID_VALUE = 1,   /* bad handle, size, extent or pointer */
    CUSTEN_ERROR_HIP = 2,             /* a HIP runtime call failed           */
    CUSTEN_ERROR_OUT_OF_MEMORY = 3,   /* staging / workspace allocation      */
    CUSTEN_ERROR_NOT_SUPPORTED = 4,   /* unsupported combination             */
    CUSTEN_ERROR_ADMISSION = 5        /* request refused by the HBM budget   */
};
/* Last error of the calling thread (cleared by custenClearError). */
int custenGetLastError(void);
const char* custenGetLastErrorString(void);
void custenClearError(void);

/* Version string of the library build. */
const char* custenVersion(void);

/* sizeof(cuSten_t) as compiled into the library (ABI check for bindings). */
size_t custenHandleSize(void);

/* Built-in device point functions, retrievable as the `func` argument of the
 * Fun Create calls by callers that cannot compile device code (C, ctypes).
 * A handle created with one of these pointers runs an inlined kernel instead
 * of an indirect call; results are bitwise identical.  Names:
 *   "x_central_difference"  devArg1X : (d[l-1] - 2 d[l] + d[l+1]) * coe[0]
 *                                      (examples/src/2d_x_np_fun.cu:49-54)
 *   "y_central_difference9" devArg1Y : sum_{i<9} coe[i] d[l + (i-4) jump]
 *                                      (examples/src/2d_y_p_fun.cu:49-62)
 *   "xy_weighted_sum"       devArg1XY: sum_{j,i} coe[j nx + i] d[l + j jump + i]
 *                                      (examples/src/2d_xy_p_fun.cu:49-68)
 *   "xy_cube_minus"         devArg1XY: sum_{j,i} coe[j nx + i] (u^3 - u), u = d[l + j jump + i]
 *                                      (cuPentCahnADI/src/cuPentCahnADI.cu:164-188)
 * Sums are fma chains from 0.0 in the order written; u^3 - u is ((u*u)*u) - u
 * unfused.  Returns NULL for an unknown name. */
double* custenGetBuiltinFunction(const char* name, int deviceNum);

/* Force the indirect-call kernel for a Fun handle even when `func` is a
 * built-in (testing and benchmarking of the drop-in path). */
void custenSetForceIndirect(cuSten_t* pt_cuSten, int force);

/* FP32-storage variants: identical semantics to the FP64 entry points with
 * float data, float weights (rounded from the double weights at Compute) and an
 * FP32 fma chain.  dir: 0 = x, 1 = y, 2 = xy; periodic: 1 = p, 0 = np. */
void custenCreate2DF32(cuSten_t* pt_cuSten, int deviceNum, int numTiles, int nx, int ny,
                       int dir, int periodic, float* dataOutput, float* dataInput,
                       double* weights, int numStenHoriz, int numStenLeft, int numStenRight,
                       int numStenVert, int numStenTop, int numStenBottom);
void custenCompute2DF32(cuSten_t* pt_cuSten, bool offload);
void custenSwap2DF32(cuSten_t* pt_cuSten, float* dataInput);
void custenDestroy2DF32(cuSten_t* pt_cuSten);

/* Stream on which Compute launches for device-resident data (default: the
 * handle's first stream).  Passing 0 selects the null stream. */
void custenSetStream(cuSten_t* pt_cuSten, hipStream_t stream);

/* Tile configuration of the stencil kernel for a handle (results do not
 * depend on it).  columnsPerLane 2 (default): rowsPerWave 4 or 8; columnsPerLane
 * 1: rowsPerWave 8 or 16.  chunksPerGroup >= 1 row chunks per workgroup.  All
 * zero: autotune on the next Compute. */
void custenSetTileConfig(cuSten_t* pt_cuSten, int rowsPerWave, int chunksPerGroup);
void custenSetTileConfigEx(cuSten_t* pt_cuSten, int rowsPerWave, int chunksPerGroup,
                           int columnsPerLane);
/* The configuration Compute uses: the caller's, else the autotuner's (all 0
 * before the first tuned Compute). */
int custenGetTileConfig(cuSten_t* pt_cuSten, int* rowsPerWave, int* chunksPerGroup,
                        int* columnsPerLane);

/* Band walker for weight stencils, FP64 and FP32 storage (and the
 * weighted-sum and cube-minus point functions, xy 3x3 and 5x5 for the
 * latter): each wave streams its own 128-column strip
 * over a band of rowsPerWave * chunksPerBand rows, with rows loaded from
 * memory directly into a ring of LDS rows, 8 rows per wave in flight ahead of
 * the computed chunk.
 * rowsPerWave 4 or 8, chunksPerBand >= 1.  Results are identical to the other
 * configurations; shapes or grids the walker does not cover use the strip
 * kernel with the same rowsPerWave and chunks.  custenSetTileConfig[Ex]
 * switches it off again. */
void custenSetBandConfig(cuSten_t* pt_cuSten, int rowsPerWave, int chunksPerBand);
/* 1 if Compute uses the band walker (set by the caller or chosen by the
 * autotuner), 0 if not, negative on error. */
int custenGetBandMode(cuSten_t* pt_cuSten);

/* WENO handles: 1 (default) squares the smoothness terms in single precision
 * as the reference does, 0 squares them in double. */
void custenSetWenoSquares(cuSten_t* pt_cuSten, int singlePrecision);

/* Fill out[0..n) with the SplitMix64 uniform stream (value i from state
 * seed + (i+1)*0x9E3779B97F4A7C15, u = (z >> 11) 2^-53, a + (b - a) u). Host. */
void custenFillUniform(double* out, size_t n, uint64_t seed, double a, double b);

#ifdef __cplusplus
}
#endif

/* Lower-case aliases. */
#define CUSTEN_ALIAS_LIST(X)                                                     \
    X(Create2DXnp) X(Swap2DXnp) X(Destroy2DXnp) X(Compute2DXnp)                  \
    X(Create2DXp) X(Swap2DXp) X(Destroy2DXp) X(Compute2DXp)                      \
    X(Create2DXnpFun) X(Swap2DXnpFun) X(Destroy2DXnpFun) X(Compute2DXnpFun)      \
    X(Create2DXpFun) X(Swap2DXpFun) X(Destroy2DXpFun) X(Compute2DXpFun)          \
    X(Create2DYnp) X(Swap2DYnp) X(Destroy2DYnp) X(Compute2DYnp)                  \
    X(Create2DYp) X(Swap2DYp) X(Destroy2DYp) X(Compute2DYp)                      \
    X(Create2DYnpFun) X(Swap2DYnpFun) X(Destroy2DYnpFun) X(Compute2DYnpFun)      \
    X(Create2DYpFun) X(Swap2DYpFun) X(Destroy2DYpFun) X(Compute2DYpFun)          \
    X(Create2DXYnp) X(Swap2DXYnp) X(Destroy2DXYnp) X(Compute2DXYnp)              \
    X(Create2DXYnpFun) X(Swap2DXYnpFun) X(Destroy2DXYnpFun) X(Compute2DXYnpFun)  \
    X(Create2DXYp) X(Swap2DXYp) X(Destroy2DXYp) X(Compute2DXYp)                  \
    X(Create2DXYpFun) X(Swap2DXYpFun) X(Destroy2DXYpFun) X(Compute2DXYpFun)      \
    X(Create2DXYWENOADVp) X(Swap2DXYWENOADVp) X(Destroy2DXYWENOADVp)             \
    X(Compute2DXYWENOADVp)

#ifndef CUSTEN_NO_LOWERCASE_ALIASES
#ifdef __cplusplus
extern "C" {
#endif
#define CUSTEN_DECLARE_ALIAS(name) __typeof__(cuSten##name) custen##name;
CUSTEN_ALIAS_LIST(CUSTEN_DECLARE_ALIAS)
#undef CUSTEN_DECLARE_ALIAS
#ifdef __cplusplus
}
#endif
#endif

#endif /* CUSTEN_H */
