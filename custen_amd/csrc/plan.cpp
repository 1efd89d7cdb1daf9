// plan.cpp -- the cuSten C API: Create / Compute / Swap / Destroy for the
// twelve x / y / xy, periodic / non-periodic, weights / point-function
// variants, the FP32-storage variant and the error utilities.
//
// Reference behaviour followed (cuSten/src/struct/custenCreateDestroy2D*.cu,
// cuSten/src/kernels/2d_*_kernel.cu):
//   Create   stores the parameters, creates 3 streams and 2 events, sets
//            nyTile = ny / numTiles, the per-tile input/output pointer tables
//            and (y, xy) the per-tile halo pointer tables
//            (custenCreateDestroy2DXYp.cu:60-241).
//   Compute  applies the stencil to the whole domain
//            (2d_xy_p_kernel.cu:542-657).
//   Swap     exchanges the per-tile input and output pointers and re-bases the
//            halo pointers on the argument (custenCreateDestroy2DXYp.cu:253-301).
//   Destroy  releases streams, events and tables; never frees user data
//            (custenCreateDestroy2DXYp.cu:312-358).
// Execution differs by design: device-resident data is computed by one
// launch over all tiles; host-resident data is streamed through HBM tile by
// tile with explicit copies on three streams (see computeStreamed).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "plan_internal.h"
#include "tuner.h"

namespace custen {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local int g_err = CUSTEN_SUCCESS;
thread_local std::string g_msg;
}  // namespace

void setError(int code, const std::string& msg)
{
    g_err = code;
    g_msg = msg;
}

void setHipError(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? CUSTEN_ERROR_OUT_OF_MEMORY : CUSTEN_ERROR_HIP;
    setError(code, std::string(what) + ": " + hipGetErrorString(e));
}

int lastError() { return g_err; }

MemKind classify(const void* p, int device)
{
    (void)device;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory reports an error
        return MemKind::Host;
    }
    switch (attr.type) {
    case hipMemoryTypeDevice:
    case hipMemoryTypeManaged:
    case hipMemoryTypeUnified:
    case hipMemoryTypeArray:
        return MemKind::Device;
    default:
        return attr.isManaged ? MemKind::Device : MemKind::Host;
    }
}

MemKind classifyCached(Plan* plan, const void* p, int device)
{
    for (int i = 0; i < 8; i++)
        if (plan->kindPtr[i] == p && p != nullptr) return plan->kind[i];
    const MemKind k = classify(p, device);
    const int slot = plan->kindNext;
    plan->kindPtr[slot] = p;
    plan->kind[slot] = k;
    plan->kindNext = (slot + 1) & 7;
    return k;
}

Plan* planOf(cuSten_t* h, const char* fn)
{
    if (h == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": null handle");
        return nullptr;
    }
    Plan* p = reinterpret_cast<Plan*>(h->plan);
    if (p == nullptr || p->magic != 0x43535450u) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 std::string(fn) + ": handle was not created (or already destroyed)");
        return nullptr;
    }
    return p;
}

// ---------------------------------------------------------------------------
// Create
// ---------------------------------------------------------------------------
namespace {

struct CreateArgs {
    const char* fn;
    Dir dir;
    bool periodic;
    bool fun;
    Dtype dtype;
    int deviceNum, numTiles, nx, ny, BX, BY;
    void* out;
    void* in;
    double* weights;  // weights (plain) or coefficients (Fun)
    int numStenHoriz, L, R, numStenVert, T, B;
    int numCoe;       // Fun only
    double* func;     // Fun only
    bool weno;        // WENO advection: u, v velocities, dx, dy spacings
    double* u;
    double* v;
    double dx, dy;
};

size_t elemSize(Dtype d) { return d == Dtype::F64 ? sizeof(double) : sizeof(float); }

double* offsetPtr(void* base, int64_t elems, Dtype d)
{
    return reinterpret_cast<double*>(static_cast<char*>(base) + elems * (int64_t)elemSize(d));
}

// Halo pointer tables of the y / xy variants, from a given input base.
// Two layouts, as the reference sets them:
//   wrapped  tile 0's top halo starts at row ny-T and the last tile's bottom
//            halo at row 0; tile t's at t*nyTile-T and (t+1)*nyTile.  Create
//            and Swap of every periodic variant, Create of 2DXYnp(Fun), and
//            Swap of every non-periodic variant
//            (custenCreateDestroy2DXYp.cu:194-228, 2DXYnp.cu:196-228,
//            2DYnp.cu:248-279, 2DXYnp.cu:265-290).
//   clamped  tile 0's top halo at row 0 and the last tile's bottom halo at
//            row ny-B: Create of 2DYnp(Fun) (custenCreateDestroy2DYnp.cu:183-214,
//            2DYnpFun.cu:186-216).
// The kernels never read these tables (halos are resolved from the whole
// grid), so the two layouts only differ in these observable fields.
void setBoundaries(cuSten_t* h, const Plan* p, void* in, bool clamped)
{
    if (h->boundaryTop == nullptr || h->boundaryBottom == nullptr) return;
    const int64_t nx = h->nx, ny = h->ny, nyT = h->nyTile;
    const int nt = h->numTiles;
    for (int t = 0; t < nt; t++) {
        int64_t top = (t == 0) ? (clamped ? 0 : ny - p->T) : (int64_t)t * nyT - p->T;
        if (top < 0) top = ((top % ny) + ny) % ny;  // halo taller than a tile
        int64_t bot = (t == nt - 1) ? (clamped ? ny - (p->V - 1 - p->T) : 0) : (int64_t)(t + 1) * nyT;
        h->boundaryTop[t] = offsetPtr(in, top * nx, p->dtype);
        h->boundaryBottom[t] = offsetPtr(in, bot * nx, p->dtype);
    }
}

void releaseHandle(cuSten_t* h, Plan* p)
{
    if (p) {
        for (int b = 0; b < 2; b++) {
            if (p->stageIn[b]) (void)hipFree(p->stageIn[b]);
            if (p->stageOut[b]) (void)hipFree(p->stageOut[b]);
            if (p->stageSide[b]) (void)hipFree(p->stageSide[b]);
            if (p->evH2D[b]) (void)hipEventDestroy(p->evH2D[b]);
            if (p->evComp[b]) (void)hipEventDestroy(p->evComp[b]);
            if (p->evD2H[b]) (void)hipEventDestroy(p->evD2H[b]);
        }
        if (p->paramDev) (void)hipFree(p->paramDev);
        p->magic = 0;
        delete p;
    }
    if (h->streams) {
        for (int s = 0; s < h->numStreams; s++)
            if (h->streams[s]) (void)hipStreamDestroy(h->streams[s]);
        std::free(h->streams);
    }
    if (h->events) {
        for (int e = 0; e < 2; e++)
            if (h->events[e]) (void)hipEventDestroy(h->events[e]);
        std::free(h->events);
    }
    std::free(h->dataInput);
    std::free(h->dataOutput);
    std::free(h->boundaryTop);
    std::free(h->boundaryBottom);
    std::free(h->uVel);
    std::free(h->vVel);
    h->streams = nullptr;
    h->events = nullptr;
    h->dataInput = h->dataOutput = h->boundaryTop = h->boundaryBottom = nullptr;
    h->uVel = h->vVel = nullptr;
    h->plan = nullptr;
}

bool validate(const CreateArgs& a, std::string& why)
{
    char buf[256];
    auto bad = [&](const char* m) {
        std::snprintf(buf, sizeof(buf), "%s: %s", a.fn, m);
        why = buf;
        return false;
    };
    if (a.nx < 1 || a.ny < 1) return bad("nx and ny must be >= 1");
    if (a.numTiles < 1) return bad("numTiles must be >= 1");
    if (a.ny % a.numTiles != 0) return bad("ny must be divisible by numTiles");
    if (a.BX < 1 || a.BY < 1) return bad("BLOCK_X and BLOCK_Y must be >= 1");
    if (a.L < 0 || a.R < 0 || a.T < 0 || a.B < 0) return bad("stencil extents must be >= 0");
    if (a.numStenHoriz != a.L + 1 + a.R)
        return bad("horizontal stencil size must equal numStenLeft + 1 + numStenRight");
    if (a.numStenVert != a.T + 1 + a.B)
        return bad("vertical stencil size must equal numStenTop + 1 + numStenBottom");
    if (a.numStenHoriz > 65 || a.numStenVert > 65) return bad("stencil extent above 65 points");
    if (a.in == nullptr || a.out == nullptr) return bad("null data pointer");
    if (a.in == a.out) return bad("input and output must not alias");
    if (a.weno) {
        if (a.u == nullptr || a.v == nullptr) return bad("null velocity pointer");
        if (a.u == a.out || a.v == a.out) return bad("velocities must not alias the output");
        if (!(std::isfinite(a.dx) && std::isfinite(a.dy)) || a.dx == 0.0 || a.dy == 0.0)
            return bad("dx and dy must be finite and nonzero");
    } else if (a.fun) {
        if (a.func == nullptr) return bad("null device function pointer");
        if (a.numCoe < 0 || a.numCoe > 1024) return bad("numCoe must be in [0, 1024]");
        if (a.numCoe > 0 && a.weights == nullptr) return bad("null coefficient pointer");
    } else if (a.weights == nullptr) {
        return bad("null weights pointer");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || a.deviceNum < 0 || a.deviceNum >= ndev)
        return bad("invalid deviceNum");
    return true;
}

Builtin recognise(const double* func, int device)
{
    if (func == nullptr) return Builtin::None;
    static const Builtin all[] = {Builtin::XCentralDiff, Builtin::YCentralDiff9,
                                  Builtin::XYWeightedSum, Builtin::XYCubeMinus};
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    Builtin found = Builtin::None;
    for (Builtin b : all) {
        void* addr = nullptr;
        if (builtinFunctionAddress(b, &addr) == hipSuccess && addr == func) {
            found = b;
            break;
        }
    }
    (void)hipGetLastError();
    (void)hipSetDevice(cur);
    return found;
}

void createCommon(cuSten_t* h, const CreateArgs& a)
{
    if (h == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(a.fn) + ": null handle");
        return;
    }
    std::memset(h, 0, sizeof(*h));
    std::string why;
    if (!validate(a, why)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, why);
        return;
    }
    hipError_t e = hipSetDevice(a.deviceNum);
    if (e != hipSuccess) {
        setHipError(e, a.fn);
        return;
    }

    Plan* p = new (std::nothrow) Plan();
    if (!p) {
        setError(CUSTEN_ERROR_OUT_OF_MEMORY, std::string(a.fn) + ": out of host memory");
        return;
    }
    p->dir = a.dir;
    p->periodic = a.periodic;
    p->fun = a.fun;
    p->dtype = a.dtype;
    p->H = a.numStenHoriz;
    p->L = a.L;
    p->V = a.numStenVert;
    p->T = a.T;
    p->numCoe = a.numCoe;
    p->zeroRight = (!a.fun && !a.periodic && a.dir == Dir::X);
    p->builtin = a.fun ? recognise(a.func, a.deviceNum) : Builtin::None;
    h->plan = reinterpret_cast<custenPlan*>(p);

    h->deviceNum = a.deviceNum;
    h->numStreams = 3;
    h->numTiles = a.numTiles;
    h->nx = a.nx;
    h->ny = a.ny;
    h->nyTile = a.ny / a.numTiles;
    h->numStenLeft = a.L;
    h->numStenRight = a.R;
    h->numStenTop = a.T;
    h->numStenBottom = a.B;
    h->numStenHoriz = a.numStenHoriz;
    h->numStenVert = a.numStenVert;
    h->numSten = (a.dir == Dir::X) ? a.numStenHoriz
                 : (a.dir == Dir::Y) ? a.numStenVert
                                     : a.numStenHoriz * a.numStenVert;
    h->BLOCK_X = a.BX;
    h->BLOCK_Y = a.BY;
    h->xGrid = (a.nx + a.BX - 1) / a.BX;
    h->yGrid = (h->nyTile + a.BY - 1) / a.BY;
    h->nxLocal = (a.dir == Dir::Y) ? a.BX : a.BX + a.L + a.R;
    h->nyLocal = (a.dir == Dir::X) ? a.BY : a.BY + a.T + a.B;
    h->mem_shared = a.weno ? wenoLdsBytes(0) : stencilLdsBytes(a.dtype, p->H, p->V, 0);
    if (a.weno) {
        // custenCreateDestroy2DXYADVWENOp.cu:99-117: coefficients 1/dx, 1/dy,
        // a 7 x 7 window (3 points each side)
        p->weno = true;
        h->coeDx = 1.0 / a.dx;
        h->coeDy = 1.0 / a.dy;
    } else if (a.fun) {
        h->coe = a.weights;
        h->numCoe = a.numCoe;
        h->devFunc = a.func;
    } else {
        h->weights = a.weights;
    }

    h->streams = static_cast<hipStream_t*>(std::calloc(3, sizeof(hipStream_t)));
    h->events = static_cast<hipEvent_t*>(std::calloc(2, sizeof(hipEvent_t)));
    h->dataInput = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
    h->dataOutput = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
    const bool halos = a.dir != Dir::X;
    if (halos) {
        h->boundaryTop = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
        h->boundaryBottom = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
    }
    if (a.weno) {
        h->uVel = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
        h->vVel = static_cast<double**>(std::calloc(a.numTiles, sizeof(double*)));
    }
    if (!h->streams || !h->events || !h->dataInput || !h->dataOutput ||
        (halos && (!h->boundaryTop || !h->boundaryBottom)) ||
        (a.weno && (!h->uVel || !h->vVel))) {
        releaseHandle(h, p);
        setError(CUSTEN_ERROR_OUT_OF_MEMORY, std::string(a.fn) + ": out of host memory");
        return;
    }
    for (int s = 0; s < 3; s++) {
        e = hipStreamCreate(&h->streams[s]);
        if (e != hipSuccess) {
            releaseHandle(h, p);
            setHipError(e, a.fn);
            return;
        }
    }
    for (int v = 0; v < 2; v++) {
        e = hipEventCreateWithFlags(&h->events[v], hipEventDisableTiming);
        if (e != hipSuccess) {
            releaseHandle(h, p);
            setHipError(e, a.fn);
            return;
        }
    }
    const int64_t tileElems = (int64_t)a.nx * h->nyTile;
    for (int t = 0; t < a.numTiles; t++) {
        h->dataInput[t] = offsetPtr(a.in, t * tileElems, a.dtype);
        h->dataOutput[t] = offsetPtr(a.out, t * tileElems, a.dtype);
        if (a.weno) {
            h->uVel[t] = a.u + t * tileElems;
            h->vVel[t] = a.v + t * tileElems;
        }
    }
    if (halos) {
        h->numBoundaryTop = a.T * a.nx;
        h->numBoundaryBottom = a.B * a.nx;
        setBoundaries(h, p, a.in, !a.periodic && a.dir == Dir::Y);
    }
}

// ---------------------------------------------------------------------------
// Compute
// ---------------------------------------------------------------------------

// Make `src` (count elements of parameter data) readable by kernels.
const double* deviceParams(Plan* p, const double* src, int count, int device, hipStream_t st,
                           const char* fn)
{
    if (src == nullptr || count <= 0) return src;
    if (classifyCached(p, src, device) == MemKind::Device) return src;
    const size_t bytes = sizeof(double) * (size_t)count;
    if (p->paramCap < bytes) {
        if (p->paramDev) (void)hipFree(p->paramDev);
        p->paramDev = nullptr;
        p->paramCap = 0;
        hipError_t e = hipMalloc(&p->paramDev, bytes);
        if (e != hipSuccess) {
            setHipError(e, fn);
            return nullptr;
        }
        p->paramCap = bytes;
    }
    hipError_t e = hipMemcpyAsync(p->paramDev, src, bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return nullptr;
    }
    return p->paramDev;
}

struct Region {
    const void* in;
    void* out;
    int64_t ldIn, ldOut;
    int nx, ny;
    RowMap rows;
    int yGlobal0;
    const double* u = nullptr;  // WENO velocities, region-local, row stride ldOut
    const double* v = nullptr;
};

hipError_t launchRegion(const cuSten_t* h, const Plan* p, const Region& r, const double* params,
                        hipStream_t st)
{
    const int R = p->H - 1 - p->L;
    const int B = p->V - 1 - p->T;
    const int wx0 = p->periodic ? 0 : p->L;
    const int wx1 = p->periodic ? h->nx : h->nx - R;
    const int wy0 = p->periodic ? 0 : p->T;
    const int wy1 = p->periodic ? h->ny : h->ny - B;
    if (p->weno) {
        WenoLaunch w{};
        w.in = static_cast<const double*>(r.in);
        w.u = r.u;
        w.v = r.v;
        w.out = static_cast<double*>(r.out);
        w.ldIn = r.ldIn;
        w.ldSide = r.ldOut;
        w.ldOut = r.ldOut;
        w.nx = r.nx;
        w.ny = r.ny;
        w.rows = r.rows;
        w.coeDx = h->coeDx;
        w.coeDy = h->coeDy;
        w.singleSquares = p->wenoSingleSquares ? 1 : 0;
        w.rowsPerWave = p->rowsPerWave;
        return launchWeno(w, st);
    }
    if (!p->fun) {
        StencilLaunch s{};
        s.dtype = p->dtype;
        s.in = r.in;
        s.out = r.out;
        s.weights = params;
        s.ldIn = r.ldIn;
        s.ldOut = r.ldOut;
        s.nx = r.nx;
        s.ny = r.ny;
        s.H = p->H;
        s.L = p->L;
        s.V = p->V;
        s.T = p->T;
        s.xWrap = p->periodic ? 1 : 0;
        s.rows = r.rows;
        s.wx0 = wx0;
        s.wx1 = wx1;
        s.wy0 = wy0;
        s.wy1 = wy1;
        s.yGlobal0 = r.yGlobal0;
        s.zeroRight = p->zeroRight ? 1 : 0;
        s.xform = Xform::Identity;
        if (p->band) {
            s.rowsPerWave = p->rowsPerWave;
            s.chunks = p->chunks;
            s.colsPerLane = 2;
            s.band = 1;
        } else {
            s.rowsPerWave = p->rowsPerWave ? p->rowsPerWave : (p->tuned ? p->tunedRpw : 0);
            s.chunks = p->chunks ? p->chunks : (p->tuned ? p->tunedChunks : 0);
            s.colsPerLane = p->colsPerLane ? p->colsPerLane : (p->tuned ? p->tunedCpl : 0);
            const bool user = p->rowsPerWave || p->chunks || p->colsPerLane;
            s.band = !user && p->tuned && p->tunedBand ? 1 : 0;
        }
        return launchStencil(s, st);
    }
    if (p->pointLauncher) {
        // the caller's kernel with its point function inlined (custen_point.h)
        custenPointArgs pa{};
        pa.in = static_cast<const double*>(r.in);
        pa.out = static_cast<double*>(r.out);
        pa.coe = params;
        pa.numCoe = p->numCoe;
        pa.dir = static_cast<int>(p->dir);
        pa.ldIn = r.ldIn;
        pa.ldOut = r.ldOut;
        pa.nx = r.nx;
        pa.ny = r.ny;
        pa.H = p->H;
        pa.L = p->L;
        pa.V = p->V;
        pa.T = p->T;
        pa.xWrap = p->periodic ? 1 : 0;
        pa.rowWrap = r.rows.wrap ? r.rows.rowWrap : 0;
        pa.rowMin = r.rows.rowMin;
        pa.rowMax = r.rows.rowMax;
        pa.wx0 = wx0;
        pa.wx1 = wx1;
        pa.wy0 = wy0;
        pa.wy1 = wy1;
        pa.yGlobal0 = r.yGlobal0;
        return p->pointLauncher(&pa, st);
    }
    FunLaunch f{};
    f.in = static_cast<const double*>(r.in);
    f.out = static_cast<double*>(r.out);
    f.coe = params;
    f.numCoe = p->numCoe;
    f.func = h->devFunc;
    f.builtin = p->forceIndirect ? Builtin::None : p->builtin;
    f.dir = static_cast<int>(p->dir);
    f.ldIn = r.ldIn;
    f.ldOut = r.ldOut;
    f.nx = r.nx;
    f.ny = r.ny;
    f.H = p->H;
    f.L = p->L;
    f.V = p->V;
    f.T = p->T;
    f.xWrap = p->periodic ? 1 : 0;
    f.rows = r.rows;
    f.wx0 = wx0;
    f.wx1 = wx1;
    f.wy0 = wy0;
    f.wy1 = wy1;
    f.yGlobal0 = r.yGlobal0;
    if (p->band) {
        f.rowsPerWave = p->rowsPerWave;
        f.chunks = p->chunks;
        f.colsPerLane = 2;
        f.band = 1;
    } else {
        f.rowsPerWave = p->rowsPerWave ? p->rowsPerWave : (p->tuned ? p->tunedRpw : 0);
        f.chunks = p->chunks ? p->chunks : (p->tuned ? p->tunedChunks : 0);
        f.colsPerLane = p->colsPerLane ? p->colsPerLane : (p->tuned ? p->tunedCpl : 0);
        const bool user = p->rowsPerWave || p->chunks || p->colsPerLane;
        f.band = !user && p->tuned && p->tunedBand ? 1 : 0;
    }
    return launchFun(f, st);
}

bool tablesConsistent(const cuSten_t* h, const Plan* p)
{
    const int64_t tileElems = (int64_t)h->nx * h->nyTile;
    for (int t = 1; t < h->numTiles; t++) {
        if (h->dataInput[t] != offsetPtr(h->dataInput[0], t * tileElems, p->dtype)) return false;
        if (h->dataOutput[t] != offsetPtr(h->dataOutput[0], t * tileElems, p->dtype))
            return false;
        if (p->weno && (h->uVel[t] != h->uVel[0] + t * tileElems ||
                        h->vVel[t] != h->vVel[0] + t * tileElems))
            return false;
    }
    return true;
}

// Host-resident data: stream tile by tile through HBM.
//   streams[1]: host -> device copies of tile t (body + halo rows)
//   streams[0]: kernel for tile t
//   streams[2]: device -> host copy of the written window of tile t
// Two staging slots alternate, so the copy-in of tile t+1 and the copy-out of
// tile t-1 overlap the kernel of tile t.
void computeStreamed(cuSten_t* h, Plan* p, const double* params, const char* fn)
{
    const size_t es = elemSize(p->dtype);
    const int64_t nx = h->nx, ny = h->ny, nyT = h->nyTile;
    const int T = p->T, B = p->V - 1 - p->T;
    const int64_t inRows = nyT + T + B;
    const size_t inBytes = (size_t)(inRows * nx) * es;
    const size_t outBytes = (size_t)(nyT * nx) * es;
    hipError_t e;
    if (!p->streamEventsReady) {
        for (int b = 0; b < 2; b++) {
            if ((e = hipEventCreateWithFlags(&p->evH2D[b], hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&p->evComp[b], hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&p->evD2H[b], hipEventDisableTiming)) != hipSuccess) {
                setHipError(e, fn);
                return;
            }
        }
        p->streamEventsReady = true;
    }
    if (p->stageInBytes < inBytes || p->stageOutBytes < outBytes) {
        for (int b = 0; b < 2; b++) {
            if (p->stageIn[b]) (void)hipFree(p->stageIn[b]);
            if (p->stageOut[b]) (void)hipFree(p->stageOut[b]);
            p->stageIn[b] = p->stageOut[b] = nullptr;
        }
        p->stageInBytes = p->stageOutBytes = 0;
        const int slots = h->numTiles > 1 ? 2 : 1;
        for (int b = 0; b < slots; b++) {
            if ((e = hipMalloc(&p->stageIn[b], inBytes)) != hipSuccess ||
                (e = hipMalloc(&p->stageOut[b], outBytes)) != hipSuccess) {
                setHipError(e, fn);
                return;
            }
        }
        p->stageInBytes = inBytes;
        p->stageOutBytes = outBytes;
    }
    // WENO: u and v of the tile's own rows (no halo) share one slot buffer
    const size_t sideBytes = p->weno ? 2 * outBytes : 0;
    if (p->stageSideBytes < sideBytes) {
        for (int b = 0; b < 2; b++) {
            if (p->stageSide[b]) (void)hipFree(p->stageSide[b]);
            p->stageSide[b] = nullptr;
        }
        p->stageSideBytes = 0;
        const int slots = h->numTiles > 1 ? 2 : 1;
        for (int b = 0; b < slots; b++) {
            if ((e = hipMalloc(&p->stageSide[b], sideBytes)) != hipSuccess) {
                setHipError(e, fn);
                return;
            }
        }
        p->stageSideBytes = sideBytes;
    }
    const char* inHost = reinterpret_cast<const char*>(h->dataInput[0]);
    char* outHost = reinterpret_cast<char*>(h->dataOutput[0]);
    hipStream_t sComp = h->streams[0], sIn = h->streams[1], sOut = h->streams[2];
    const int R = p->H - 1 - p->L;
    // written window (np variants leave the border of the host array untouched)
    int64_t wxLo = 0, wxHi = nx;
    if (!p->periodic) {
        wxLo = p->L;
        wxHi = p->zeroRight ? nx : nx - R;
    }
    const int64_t wyLo = p->periodic ? 0 : T;
    const int64_t wyHi = p->periodic ? ny : ny - B;
    const int slots = h->numTiles > 1 ? 2 : 1;

    for (int t = 0; t < h->numTiles; t++) {
        const int b = t % slots;
        char* sin = static_cast<char*>(p->stageIn[b]);
        char* sout = static_cast<char*>(p->stageOut[b]);
        // slot b is free once the kernel of tile t-2 finished reading it
        if (t >= slots) {
            if ((e = hipStreamWaitEvent(sIn, p->evComp[b], 0)) != hipSuccess) break;
        }
        // copy rows [t*nyT - T, (t+1)*nyT + B) (wrapped or clamped) in maximal runs
        const int64_t g0 = (int64_t)t * nyT - T;
        int64_t k = 0;
        e = hipSuccess;
        while (k < inRows && e == hipSuccess) {
            int64_t gy = g0 + k;
            int64_t src;
            int64_t run;
            if (p->periodic) {
                src = ((gy % ny) + ny) % ny;
                run = std::min<int64_t>(inRows - k, ny - src);
            } else if (gy < 0 || gy >= ny) {
                src = gy < 0 ? 0 : ny - 1;  // never read by a written point
                run = 1;
            } else {
                src = gy;
                run = std::min<int64_t>(inRows - k, ny - gy);
            }
            e = hipMemcpyAsync(sin + (size_t)(k * nx) * es, inHost + (size_t)(src * nx) * es,
                               (size_t)(run * nx) * es, hipMemcpyDefault, sIn);
            k += run;
        }
        if (e != hipSuccess) break;
        if (p->weno) {
            char* side = static_cast<char*>(p->stageSide[b]);
            const size_t off = (size_t)((int64_t)t * nyT * nx) * es;
            if ((e = hipMemcpyAsync(side, reinterpret_cast<const char*>(h->uVel[0]) + off,
                                    outBytes, hipMemcpyDefault, sIn)) != hipSuccess ||
                (e = hipMemcpyAsync(side + outBytes, reinterpret_cast<const char*>(h->vVel[0]) + off,
                                    outBytes, hipMemcpyDefault, sIn)) != hipSuccess)
                break;
        }
        if ((e = hipEventRecord(p->evH2D[b], sIn)) != hipSuccess) break;
        if ((e = hipStreamWaitEvent(sComp, p->evH2D[b], 0)) != hipSuccess) break;
        if (t >= slots) {
            if ((e = hipStreamWaitEvent(sComp, p->evD2H[b], 0)) != hipSuccess) break;
        }
        Region r;
        r.in = sin + (size_t)(T * nx) * es;
        r.out = sout;
        r.ldIn = nx;
        r.ldOut = nx;
        r.nx = (int)nx;
        r.ny = (int)nyT;
        r.rows.wrap = 0;
        r.rows.rowWrap = 0;
        r.rows.rowMin = -T;
        r.rows.rowMax = (int)(nyT + B - 1);
        r.yGlobal0 = (int)((int64_t)t * nyT);
        if (p->weno) {
            r.u = static_cast<const double*>(p->stageSide[b]);
            r.v = r.u + nyT * nx;
        }
        if ((e = launchRegion(h, p, r, params, sComp)) != hipSuccess) break;
        if ((e = hipEventRecord(p->evComp[b], sComp)) != hipSuccess) break;
        if ((e = hipStreamWaitEvent(sOut, p->evComp[b], 0)) != hipSuccess) break;
        const int64_t y0 = std::max<int64_t>((int64_t)t * nyT, wyLo);
        const int64_t y1 = std::min<int64_t>((int64_t)(t + 1) * nyT, wyHi);
        if (y1 > y0 && wxHi > wxLo) {
            const int64_t ly0 = y0 - (int64_t)t * nyT;
            if (wxLo == 0 && wxHi == nx) {
                e = hipMemcpyAsync(outHost + (size_t)(y0 * nx) * es, sout + (size_t)(ly0 * nx) * es,
                                   (size_t)((y1 - y0) * nx) * es, hipMemcpyDefault, sOut);
            } else {
                e = hipMemcpy2DAsync(outHost + (size_t)(y0 * nx + wxLo) * es, (size_t)nx * es,
                                     sout + (size_t)(ly0 * nx + wxLo) * es, (size_t)nx * es,
                                     (size_t)(wxHi - wxLo) * es, (size_t)(y1 - y0),
                                     hipMemcpyDefault, sOut);
            }
            if (e != hipSuccess) break;
        }
        if ((e = hipEventRecord(p->evD2H[b], sOut)) != hipSuccess) break;
    }
    if (e != hipSuccess) {
        setHipError(e, fn);
        return;
    }
    e = hipStreamSynchronize(sOut);
    if (e != hipSuccess) setHipError(e, fn);
}

void computeCommon(cuSten_t* h, bool offload, const char* fn)
{
    (void)offload;  // placement hint only: results are valid on the host either way
    Plan* p = planOf(h, fn);
    if (!p) return;
    hipError_t e = hipSetDevice(h->deviceNum);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return;
    }
    if (!tablesConsistent(h, p)) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 std::string(fn) + ": per-tile pointer tables are not contiguous");
        return;
    }
    const void* in = h->dataInput[0];
    void* out = h->dataOutput[0];
    const int dev = h->deviceNum;
    const bool resident =
        classifyCached(p, in, dev) == MemKind::Device &&
        classifyCached(p, out, dev) == MemKind::Device &&
        (!p->weno || (classifyCached(p, h->uVel[0], dev) == MemKind::Device &&
                      classifyCached(p, h->vVel[0], dev) == MemKind::Device));
    hipStream_t st = p->hasUserStream ? p->userStream : h->streams[0];
    const double* src = p->weno ? nullptr : p->fun ? h->coe : h->weights;
    const int count = p->weno ? 0 : p->fun ? p->numCoe : p->H * p->V;
    const double* params = deviceParams(p, src, count, h->deviceNum,
                                        resident ? st : h->streams[1], fn);
    if (src != nullptr && count > 0 && params == nullptr) return;
    if (resident) {
        Region r;
        r.in = in;
        r.out = out;
        r.ldIn = h->nx;
        r.ldOut = h->nx;
        r.nx = h->nx;
        r.ny = h->ny;
        if (p->periodic) {
            r.rows.wrap = 1;
            r.rows.rowWrap = h->ny;
            r.rows.rowMin = 0;
            r.rows.rowMax = h->ny - 1;
        } else {
            r.rows.wrap = 0;
            r.rows.rowWrap = 0;
            r.rows.rowMin = 0;
            r.rows.rowMax = h->ny - 1;
        }
        r.yGlobal0 = 0;
        if (p->weno) {
            r.u = h->uVel[0];
            r.v = h->vVel[0];
        }
        // First resident Compute of a large grid: pick the tile configuration
        // (bit-identical results for every candidate; synchronous once).
        if (!p->weno && !p->tuned && !p->band && p->rowsPerWave == 0 && p->chunks == 0 &&
            p->colsPerLane == 0 &&
            tunerEnabled() &&
            (int64_t)h->nx * h->ny >= (int64_t(1) << 20) && !streamCapturing(st)) {
            TuneKey k;
            k.kind = 0;
            k.dtype = static_cast<int>(p->dtype);
            k.H = p->H;
            k.V = p->V;
            k.xform = p->fun ? 100 + static_cast<int>(p->forceIndirect ? Builtin::None : p->builtin)
                             : 0;
            k.nx = h->nx;
            k.ny = h->ny;
            TuneVal v;
            if (!tunerLookup(k, v)) {
                // one-column-per-lane strips (custenSetTileConfigEx) measured
                // slower on every shape, so they are not candidates
                std::vector<TuneVal> cands = {{4, 1, 2}, {4, 2, 2}, {4, 4, 2}, {8, 1, 2},
                                              {8, 2, 2}, {8, 4, 2}, {8, 8, 2}, {8, 16, 2}};
                // band-walker candidates (rows per wave, chunks per band):
                // weight stencils and the weighted-sum and cube-minus point
                // functions (both run on the stencil kernels)
                const bool bandOk =
                    !p->fun || (p->dir == Dir::XY && !p->forceIndirect &&
                                (p->builtin == Builtin::XYWeightedSum ||
                                 p->builtin == Builtin::XYCubeMinus));
                if (bandOk) {
                    for (const TuneVal& b :
                         {TuneVal{4, 2, 2, 1}, TuneVal{4, 4, 2, 1}, TuneVal{4, 8, 2, 1},
                          TuneVal{8, 1, 2, 1}, TuneVal{8, 2, 2, 1}, TuneVal{8, 4, 2, 1}})
                        cands.push_back(b);
                }
                auto run = [&](const TuneVal& c) {
                    Plan tmp = *p;
                    tmp.rowsPerWave = c.rpw;
                    tmp.chunks = c.chunks;
                    tmp.colsPerLane = c.cpl;
                    tmp.band = c.band != 0;
                    return launchRegion(h, &tmp, r, params, st);
                };
                e = tunerMeasure(st, run, cands, 4, v);
                if (e != hipSuccess) {
                    setHipError(e, fn);
                    return;
                }
                tunerStore(k, v);
            }
            p->tuned = true;
            p->tunedRpw = v.rpw;
            p->tunedChunks = v.chunks;
            p->tunedCpl = v.cpl;
            p->tunedBand = v.band != 0;
        }
        e = launchRegion(h, p, r, params, st);
        if (e != hipSuccess) setHipError(e, fn);
        return;
    }
    if (p->hasUserStream) {
        // order the streaming pipeline after prior work on the user's stream
        if ((e = hipEventRecord(h->events[0], st)) != hipSuccess ||
            (e = hipStreamWaitEvent(h->streams[1], h->events[0], 0)) != hipSuccess) {
            setHipError(e, fn);
            return;
        }
    }
    computeStreamed(h, p, params, fn);
}

void swapCommon(cuSten_t* h, void* dataInput, const char* fn)
{
    Plan* p = planOf(h, fn);
    if (!p) return;
    for (int t = 0; t < h->numTiles; t++) std::swap(h->dataInput[t], h->dataOutput[t]);
    if (dataInput == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": null dataInput");
        return;
    }
    if (dataInput != h->dataInput[0])
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 std::string(fn) + ": dataInput is not the swapped input base; halo pointers "
                                   "follow the argument, tiles follow the swap");
    setBoundaries(h, p, dataInput, false);
}

void destroyCommon(cuSten_t* h, const char* fn)
{
    Plan* p = planOf(h, fn);
    if (!p) return;
    (void)hipSetDevice(h->deviceNum);
    for (int s = 0; s < h->numStreams; s++)
        if (h->streams && h->streams[s]) (void)hipStreamSynchronize(h->streams[s]);
    releaseHandle(h, p);
}

CreateArgs baseArgs(const char* fn, Dir dir, bool periodic, bool fun, int deviceNum, int numTiles,
                    int nx, int ny, int BX, int BY, void* out, void* in, double* w)
{
    CreateArgs a{};
    a.fn = fn;
    a.dir = dir;
    a.periodic = periodic;
    a.fun = fun;
    a.dtype = Dtype::F64;
    a.deviceNum = deviceNum;
    a.numTiles = numTiles;
    a.nx = nx;
    a.ny = ny;
    a.BX = BX;
    a.BY = BY;
    a.out = out;
    a.in = in;
    a.weights = w;
    a.numStenHoriz = 1;
    a.numStenVert = 1;
    return a;
}

void createX(const char* fn, bool periodic, cuSten_t* h, int dev, int nt, int nx, int ny, int BX,
             int BY, double* out, double* in, double* w, int numSten, int L, int R)
{
    CreateArgs a = baseArgs(fn, Dir::X, periodic, false, dev, nt, nx, ny, BX, BY, out, in, w);
    a.numStenHoriz = numSten;
    a.L = L;
    a.R = R;
    createCommon(h, a);
}

void createXFun(const char* fn, bool periodic, cuSten_t* h, int dev, int nt, int nx, int ny,
                int BX, int BY, double* out, double* in, double* coe, int numSten, int L, int R,
                int numCoe, double* func)
{
    CreateArgs a = baseArgs(fn, Dir::X, periodic, true, dev, nt, nx, ny, BX, BY, out, in, coe);
    a.numStenHoriz = numSten;
    a.L = L;
    a.R = R;
    a.numCoe = numCoe;
    a.func = func;
    createCommon(h, a);
}

void createY(const char* fn, bool periodic, cuSten_t* h, int dev, int nt, int nx, int ny, int BX,
             int BY, double* out, double* in, double* w, int numSten, int T, int B)
{
    CreateArgs a = baseArgs(fn, Dir::Y, periodic, false, dev, nt, nx, ny, BX, BY, out, in, w);
    a.numStenVert = numSten;
    a.T = T;
    a.B = B;
    createCommon(h, a);
}

void createYFun(const char* fn, bool periodic, cuSten_t* h, int dev, int nt, int nx, int ny,
                int BX, int BY, double* out, double* in, double* coe, int numSten, int T, int B,
                int numCoe, double* func)
{
    CreateArgs a = baseArgs(fn, Dir::Y, periodic, true, dev, nt, nx, ny, BX, BY, out, in, coe);
    a.numStenVert = numSten;
    a.T = T;
    a.B = B;
    a.numCoe = numCoe;
    a.func = func;
    createCommon(h, a);
}

void createXY(const char* fn, bool periodic, bool fun, cuSten_t* h, int dev, int nt, int nx,
              int ny, int BX, int BY, double* out, double* in, double* w, int Hs, int L, int R,
              int Vs, int T, int B, double* func)
{
    CreateArgs a = baseArgs(fn, Dir::XY, periodic, fun, dev, nt, nx, ny, BX, BY, out, in, w);
    a.numStenHoriz = Hs;
    a.L = L;
    a.R = R;
    a.numStenVert = Vs;
    a.T = T;
    a.B = B;
    if (fun) {
        // the xy Fun kernels stage numStenHoriz * numStenVert coefficients
        // (2d_xy_p_fun_kernel.cu:116-119)
        a.numCoe = Hs * Vs;
        a.func = func;
    }
    createCommon(h, a);
}

}  // namespace
}  // namespace custen

using namespace custen;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

// ---- x, weights -------------------------------------------------------------
void cuStenCreate2DXnp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                       int BLOCK_Y, double* dataOutput, double* dataInput, double* weights,
                       int numSten, int numStenLeft, int numStenRight)
{
    createX("cuStenCreate2DXnp", false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
            dataOutput, dataInput, weights, numSten, numStenLeft, numStenRight);
}
void cuStenSwap2DXnp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXnp"); }
void cuStenDestroy2DXnp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXnp"); }
void cuStenCompute2DXnp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXnp"); }

void cuStenCreate2DXp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                      int BLOCK_Y, double* dataNew, double* dataOld, double* weights, int numSten,
                      int numStenLeft, int numStenRight)
{
    createX("cuStenCreate2DXp", true, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y, dataNew,
            dataOld, weights, numSten, numStenLeft, numStenRight);
}
void cuStenSwap2DXp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXp"); }
void cuStenDestroy2DXp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXp"); }
void cuStenCompute2DXp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXp"); }

// ---- x, point function ------------------------------------------------------
void cuStenCreate2DXnpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                          int BLOCK_Y, double* dateOutput, double* dateInput, double* coe,
                          int numSten, int numStenLeft, int numStenRight, int numCoe, double* func)
{
    createXFun("cuStenCreate2DXnpFun", false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
               dateOutput, dateInput, coe, numSten, numStenLeft, numStenRight, numCoe, func);
}
void cuStenSwap2DXnpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXnpFun"); }
void cuStenDestroy2DXnpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXnpFun"); }
void cuStenCompute2DXnpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXnpFun"); }

void cuStenCreate2DXpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                         int BLOCK_Y, double* dataNew, double* dataOld, double* coe, int numSten,
                         int numStenLeft, int numStenRight, int numCoe, double* func)
{
    createXFun("cuStenCreate2DXpFun", true, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
               dataNew, dataOld, coe, numSten, numStenLeft, numStenRight, numCoe, func);
}
void cuStenSwap2DXpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXpFun"); }
void cuStenDestroy2DXpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXpFun"); }
void cuStenCompute2DXpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXpFun"); }
void cuSenCompute2DXpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuSenCompute2DXpFun"); }

// ---- y ----------------------------------------------------------------------
void cuStenCreate2DYnp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                       int BLOCK_Y, double* dataOutput, double* dataInput, double* weights,
                       int numSten, int numStenTop, int numStenBottom)
{
    createY("cuStenCreate2DYnp", false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
            dataOutput, dataInput, weights, numSten, numStenTop, numStenBottom);
}
void cuStenSwap2DYnp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DYnp"); }
void cuStenDestroy2DYnp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DYnp"); }
void cuStenCompute2DYnp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DYnp"); }

void cuStenCreate2DYp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                      int BLOCK_Y, double* dataOutput, double* dataInput, double* weights,
                      int numSten, int numStenTop, int numStenBottom)
{
    createY("cuStenCreate2DYp", true, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
            dataOutput, dataInput, weights, numSten, numStenTop, numStenBottom);
}
void cuStenSwap2DYp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DYp"); }
void cuStenDestroy2DYp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DYp"); }
void cuStenCompute2DYp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DYp"); }

void cuStenCreate2DYnpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                          int BLOCK_Y, double* dataOutput, double* dataInput, double* coe,
                          int numSten, int numStenTop, int numStenBottom, double* func)
{
    // 2d_y_np_fun_kernel.cu:102 stages numSten coefficients
    createYFun("cuStenCreate2DYnpFun", false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
               dataOutput, dataInput, coe, numSten, numStenTop, numStenBottom, numSten, func);
}
void cuStenSwap2DYnpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DYnpFun"); }
void cuStenDestroy2DYnpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DYnpFun"); }
void cuStenCompute2DYnpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DYnpFun"); }

void cuStenCreate2DYpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                         int BLOCK_Y, double* dataOutput, double* dataInput, double* coe,
                         int numSten, int numStenTop, int numStenBottom, int numCoe, double* func)
{
    createYFun("cuStenCreate2DYpFun", true, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
               dataOutput, dataInput, coe, numSten, numStenTop, numStenBottom, numCoe, func);
}
void cuStenSwap2DYpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DYpFun"); }
void cuStenDestroy2DYpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DYpFun"); }
void cuStenCompute2DYpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DYpFun"); }

// ---- xy -------------------------------------------------------------------
void cuStenCreate2DXYnp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                        int BLOCK_Y, double* dataOutput, double* dataInput, double* weights,
                        int numStenHoriz, int numStenLeft, int numStenRight, int numStenVert,
                        int numStenTop, int numStenBottom)
{
    createXY("cuStenCreate2DXYnp", false, false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
             dataOutput, dataInput, weights, numStenHoriz, numStenLeft, numStenRight, numStenVert,
             numStenTop, numStenBottom, nullptr);
}
void cuStenSwap2DXYnp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXYnp"); }
void cuStenDestroy2DXYnp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXYnp"); }
void cuStenCompute2DXYnp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXYnp"); }

void cuStenCreate2DXYnpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                           int BLOCK_Y, double* dataNew, double* dataOld, double* coe,
                           int numStenHoriz, int numStenLeft, int numStenRight, int numStenVert,
                           int numStenTop, int numStenBottom, double* func)
{
    createXY("cuStenCreate2DXYnpFun", false, true, h, deviceNum, numTiles, nx, ny, BLOCK_X,
             BLOCK_Y, dataNew, dataOld, coe, numStenHoriz, numStenLeft, numStenRight, numStenVert,
             numStenTop, numStenBottom, func);
}
void cuStenSwap2DXYnpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXYnpFun"); }
void cuStenDestroy2DXYnpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXYnpFun"); }
void cuStenCompute2DXYnpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXYnpFun"); }

void cuStenCreate2DXYp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                       int BLOCK_Y, double* dataNew, double* dataOld, double* weights,
                       int numStenHoriz, int numStenLeft, int numStenRight, int numStenVert,
                       int numStenTop, int numStenBottom)
{
    createXY("cuStenCreate2DXYp", true, false, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
             dataNew, dataOld, weights, numStenHoriz, numStenLeft, numStenRight, numStenVert,
             numStenTop, numStenBottom, nullptr);
}
void cuStenSwap2DXYp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXYp"); }
void cuStenDestroy2DXYp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXYp"); }
void cuStenCompute2DXYp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXYp"); }

void cuStenCreate2DXYpFun(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int BLOCK_X,
                          int BLOCK_Y, double* dataOutput, double* dataInput, double* coe,
                          int numStenHoriz, int numStenLeft, int numStenRight, int numStenVert,
                          int numStenTop, int numStenBottom, double* func)
{
    createXY("cuStenCreate2DXYpFun", true, true, h, deviceNum, numTiles, nx, ny, BLOCK_X, BLOCK_Y,
             dataOutput, dataInput, coe, numStenHoriz, numStenLeft, numStenRight, numStenVert,
             numStenTop, numStenBottom, func);
}
void cuStenSwap2DXYpFun(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXYpFun"); }
void cuStenDestroy2DXYpFun(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXYpFun"); }
void cuStenCompute2DXYpFun(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXYpFun"); }

// ---- WENO5 advection (custenCreateDestroy2DXYADVWENOp.cu:57-260) ---------------
void cuStenCreate2DXYWENOADVp(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny,
                              int BLOCK_X, int BLOCK_Y, double dx, double dy, double* u, double* v,
                              double* dataOutput, double* dataInput)
{
    CreateArgs a = baseArgs("cuStenCreate2DXYWENOADVp", Dir::XY, true, false, deviceNum, numTiles,
                            nx, ny, BLOCK_X, BLOCK_Y, dataOutput, dataInput, nullptr);
    a.numStenHoriz = 7;
    a.L = 3;
    a.R = 3;
    a.numStenVert = 7;
    a.T = 3;
    a.B = 3;
    a.weno = true;
    a.u = u;
    a.v = v;
    a.dx = dx;
    a.dy = dy;
    createCommon(h, a);
}
void cuStenSwap2DXYWENOADVp(cuSten_t* h, double* dataInput) { swapCommon(h, dataInput, "cuStenSwap2DXYWENOADVp"); }
void cuStenDestroy2DXYWENOADVp(cuSten_t* h) { destroyCommon(h, "cuStenDestroy2DXYWENOADVp"); }
void cuStenCompute2DXYWENOADVp(cuSten_t* h, bool offload) { computeCommon(h, offload, "cuStenCompute2DXYWENOADVp"); }

void custenSetWenoSquares(cuSten_t* h, int singlePrecision)
{
    Plan* p = planOf(h, "custenSetWenoSquares");
    if (!p) return;
    if (!p->weno) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenSetWenoSquares: not a WENO handle");
        return;
    }
    p->wenoSingleSquares = singlePrecision != 0;
}

// ---- FP32 storage -----------------------------------------------------------
void custenCreate2DF32(cuSten_t* h, int deviceNum, int numTiles, int nx, int ny, int dir,
                       int periodic, float* dataOutput, float* dataInput, double* weights,
                       int numStenHoriz, int numStenLeft, int numStenRight, int numStenVert,
                       int numStenTop, int numStenBottom)
{
    if (dir < 0 || dir > 2) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCreate2DF32: dir must be 0, 1 or 2");
        return;
    }
    CreateArgs a = baseArgs("custenCreate2DF32", static_cast<Dir>(dir), periodic != 0, false,
                            deviceNum, numTiles, nx, ny, 32, 32, dataOutput, dataInput, weights);
    a.dtype = Dtype::F32;
    a.numStenHoriz = dir == 1 ? 1 : numStenHoriz;
    a.L = dir == 1 ? 0 : numStenLeft;
    a.R = dir == 1 ? 0 : numStenRight;
    a.numStenVert = dir == 0 ? 1 : numStenVert;
    a.T = dir == 0 ? 0 : numStenTop;
    a.B = dir == 0 ? 0 : numStenBottom;
    createCommon(h, a);
}
void custenCompute2DF32(cuSten_t* h, bool offload) { computeCommon(h, offload, "custenCompute2DF32"); }
void custenSwap2DF32(cuSten_t* h, float* dataInput) { swapCommon(h, dataInput, "custenSwap2DF32"); }
void custenDestroy2DF32(cuSten_t* h) { destroyCommon(h, "custenDestroy2DF32"); }

// ---- extensions -------------------------------------------------------------
void custenSetStream(cuSten_t* h, hipStream_t stream)
{
    Plan* p = planOf(h, "custenSetStream");
    if (!p) return;
    p->userStream = stream;
    p->hasUserStream = true;
}

void custenSetTileConfigEx(cuSten_t* h, int rowsPerWave, int chunksPerGroup, int columnsPerLane)
{
    Plan* p = planOf(h, "custenSetTileConfigEx");
    if (!p) return;
    if (columnsPerLane != 0 && columnsPerLane != 1 && columnsPerLane != 2) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenSetTileConfigEx: columnsPerLane must be 0, 1 or 2");
        return;
    }
    const bool ok = rowsPerWave == 0 ||
                    (columnsPerLane == 1 ? (rowsPerWave == 8 || rowsPerWave == 16)
                                         : (rowsPerWave == 4 || rowsPerWave == 8));
    if (!ok) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 "custenSetTileConfigEx: rowsPerWave must be 4 or 8 (2 columns per lane), "
                 "8 or 16 (1 column per lane), or 0");
        return;
    }
    if (chunksPerGroup < 0) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenSetTileConfigEx: chunksPerGroup must be >= 0");
        return;
    }
    p->rowsPerWave = rowsPerWave;
    p->chunks = chunksPerGroup;
    p->colsPerLane = columnsPerLane;
    p->band = false;
    if (rowsPerWave == 0 && chunksPerGroup == 0 && columnsPerLane == 0)
        p->tuned = false;  // re-tune on next Compute
    h->mem_shared = stencilLdsBytes(p->dtype, p->H, p->V, rowsPerWave);
}

void custenSetBandConfig(cuSten_t* h, int rowsPerWave, int chunksPerBand)
{
    Plan* p = planOf(h, "custenSetBandConfig");
    if (!p) return;
    if (rowsPerWave != 4 && rowsPerWave != 8) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenSetBandConfig: rowsPerWave must be 4 or 8");
        return;
    }
    if (chunksPerBand < 1) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenSetBandConfig: chunksPerBand must be >= 1");
        return;
    }
    p->rowsPerWave = rowsPerWave;
    p->chunks = chunksPerBand;
    p->colsPerLane = 2;
    p->band = true;
    h->mem_shared = stencilBandLdsBytes(p->dtype, p->H, p->V, rowsPerWave);
}

int custenGetBandMode(cuSten_t* h)
{
    Plan* p = planOf(h, "custenGetBandMode");
    if (!p) return -CUSTEN_ERROR_INVALID_VALUE;
    const bool user = p->rowsPerWave || p->chunks || p->colsPerLane;
    return (p->band || (!user && p->tuned && p->tunedBand)) ? 1 : 0;
}

void custenSetTileConfig(cuSten_t* h, int rowsPerWave, int chunksPerGroup)
{
    custenSetTileConfigEx(h, rowsPerWave, chunksPerGroup, 0);
}

int custenGetTileConfig(cuSten_t* h, int* rowsPerWave, int* chunksPerGroup, int* columnsPerLane)
{
    Plan* p = planOf(h, "custenGetTileConfig");
    if (!p) return CUSTEN_ERROR_INVALID_VALUE;
    const bool user = p->rowsPerWave || p->chunks || p->colsPerLane;
    if (rowsPerWave) *rowsPerWave = user ? p->rowsPerWave : (p->tuned ? p->tunedRpw : 0);
    if (chunksPerGroup) *chunksPerGroup = user ? p->chunks : (p->tuned ? p->tunedChunks : 0);
    if (columnsPerLane) *columnsPerLane = user ? p->colsPerLane : (p->tuned ? p->tunedCpl : 0);
    return CUSTEN_SUCCESS;
}

void custenSetForceIndirect(cuSten_t* h, int force)
{
    Plan* p = planOf(h, "custenSetForceIndirect");
    if (!p) return;
    p->forceIndirect = force != 0;
}

void custenSetPointLauncher(cuSten_t* h, custenPointLauncher launcher)
{
    Plan* p = planOf(h, "custenSetPointLauncher");
    if (!p) return;
    if (!p->fun || p->weno) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 "custenSetPointLauncher: the handle is not a point-function handle");
        return;
    }
    p->pointLauncher = launcher;
}

double* custenGetBuiltinFunction(const char* name, int deviceNum)
{
    if (name == nullptr) return nullptr;
    Builtin b = Builtin::None;
    if (std::strcmp(name, "x_central_difference") == 0) b = Builtin::XCentralDiff;
    else if (std::strcmp(name, "y_central_difference9") == 0) b = Builtin::YCentralDiff9;
    else if (std::strcmp(name, "xy_weighted_sum") == 0) b = Builtin::XYWeightedSum;
    else if (std::strcmp(name, "xy_cube_minus") == 0) b = Builtin::XYCubeMinus;
    if (b == Builtin::None) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string("custenGetBuiltinFunction: unknown name ") + name);
        return nullptr;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    hipError_t e = hipSetDevice(deviceNum);
    void* addr = nullptr;
    if (e == hipSuccess) e = builtinFunctionAddress(b, &addr);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        setHipError(e, "custenGetBuiltinFunction");
        return nullptr;
    }
    return static_cast<double*>(addr);
}

int custenGetLastError(void) { return lastError(); }
const char* custenGetLastErrorString(void) { return g_msg.c_str(); }
void custenClearError(void)
{
    g_err = CUSTEN_SUCCESS;
    g_msg.clear();
}
const char* custenVersion(void) { return "custen-amd 0.1.0 (gfx950)"; }
size_t custenHandleSize(void) { return sizeof(cuSten_t); }

// util.h:43 / error.cu:41-50: on a pending HIP error, print
// "\nError while '<action>': <error string>\nprogram terminated ...\n\n" to
// stdout and exit(EXIT_FAILURE).  An error this library recorded (a refused
// Create, a failed Compute) is reported and ends the process the same way;
// the reference had no such errors because it validated nothing.
void checkError(const char* action)
{
    const char* what = action ? action : "";
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::printf("\nError while '%s': %s\nprogram terminated ...\n\n", what,
                    hipGetErrorString(e));
        std::fflush(stdout);
        std::exit(EXIT_FAILURE);
    }
    if (g_err != CUSTEN_SUCCESS) {
        std::printf("\nError while '%s': %s\nprogram terminated ...\n\n", what, g_msg.c_str());
        std::fflush(stdout);
        std::exit(EXIT_FAILURE);
    }
}

}  // extern "C"

// ---- lower-case aliases ------------------------------------------------------
// custenCreate2DXYp etc. are ELF aliases of the cuSten* symbols (same address).
#define CUSTEN_DEFINE_ALIAS(name)                                   \
    __asm__(".globl custen" #name "\n"                              \
            ".type custen" #name ", @function\n"                    \
            ".set custen" #name ", cuSten" #name "\n");
CUSTEN_ALIAS_LIST(CUSTEN_DEFINE_ALIAS)
#undef CUSTEN_DEFINE_ALIAS
