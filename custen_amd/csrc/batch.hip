// batch.hip -- many independent periodic xy grids of uneven sizes advanced
// together: one launch per step covers every strip of every grid (the
// workgroup finds its grid by binary search over the strip prefix sums), and
// a whole multi-step run is captured once into a HIP graph and replayed.
// The per-grid launch loop (one launch per grid per step) is kept as the
// baseline mode.  Per grid the arithmetic is the single-grid kernel's
// (stencil_body.h), so results are bit-identical to cuStenCompute2DXYp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/cuSten.h"
#include "../../include/custen_batch.h"
#include "kernels.h"
#include "plan_internal.h"
#include "stencil_body.h"
#include "tuner.h"

#pragma clang fp contract(off)

namespace custen {

struct BatchDev {
    const void* in;
    void* out;
    int n;          // nx == ny == n for square grids; nx, ny kept separately below
    int nx, ny;
    int stripBase;  // first strip id of this grid
    int vec;        // rows pair-aligned
};

namespace dev {

__device__ __forceinline__ int findGrid(const BatchDev* g, int count, int item)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].stripBase <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, int H, int V, int RPW>
__global__ __launch_bounds__(kThreads) void batchKernel(const BatchDev* __restrict__ grids,
                                                        int count, int firstItem, int parity,
                                                        const double* __restrict__ weights, int L,
                                                        int Tt, int chunks)
{
    __shared__ __attribute__((aligned(16))) T tile[(kWaves * RPW + V - 1) * ldsRowLength(H)];
    T w[H * V];
#pragma unroll
    for (int k = 0; k < H * V; k++) w[k] = static_cast<T>(weights[k]);
    const int item = firstItem + xcdRemap(blockIdx.x, gridDim.x);
    const int gi = findGrid(grids, count, item);
    const BatchDev& G = grids[gi];
    StencilLaunch a;
    a.dtype = sizeof(T) == 8 ? Dtype::F64 : Dtype::F32;
    a.in = parity ? G.out : G.in;
    a.out = parity ? const_cast<void*>(G.in) : G.out;
    a.weights = weights;
    a.ldIn = G.nx;
    a.ldOut = G.nx;
    a.nx = G.nx;
    a.ny = G.ny;
    a.H = H;
    a.L = L;
    a.V = V;
    a.T = Tt;
    a.xWrap = 1;
    a.rows.wrap = 1;
    a.rows.rowWrap = G.ny;
    a.rows.rowMin = 0;
    a.rows.rowMax = G.ny - 1;
    a.wx0 = 0;
    a.wx1 = G.nx;
    a.wy0 = 0;
    a.wy1 = G.ny;
    a.yGlobal0 = 0;
    a.zeroRight = 0;
    a.xform = Xform::Identity;
    a.rowsPerWave = RPW;
    a.chunks = chunks;
    a.vecIn = G.vec;
    a.vecOut = G.vec;
    stencilStrip<T, H, V, RPW, Xform::Identity>(a, w, tile, item - G.stripBase);
}

}  // namespace dev
}  // namespace custen

using namespace custen;

struct custenBatch {
    int device = 0;
    int numGrids = 0;
    Dtype dtype = Dtype::F64;
    int H = 5, L = 2, V = 5, T = 2;
    int rpw = 8, chunks = 4;
    std::vector<BatchDev> host;
    std::vector<int> strips;  // strips per grid
    int totalStrips = 0;
    BatchDev* dGrids = nullptr;
    double* dWeights = nullptr;
    int parity = 0;           // 1: the current input is the original output side
    bool tuned = false;
    hipStream_t own = nullptr;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (steps, start parity)
    std::map<std::pair<int, int>, hipGraph_t> graphDefs;
};

namespace {

int stripsFor(int nx, int ny, int rpw, int chunks)
{
    const int tilesX = (nx + dev::kTW - 1) / dev::kTW;
    const int rowsPer = dev::kWaves * rpw * chunks;
    return tilesX * ((ny + rowsPer - 1) / rowsPer);
}

template <typename T, int H, int V, int RPW>
hipError_t launchOne(const custenBatch* b, int count, const BatchDev* grids, int firstItem,
                     int nItems, int parity, hipStream_t st)
{
    hipLaunchKernelGGL((dev::batchKernel<T, H, V, RPW>), dim3(nItems), dim3(dev::kThreads), 0, st,
                       grids, count, firstItem, parity, b->dWeights, b->L, b->T, b->chunks);
    return hipGetLastError();
}

hipError_t launchItems(const custenBatch* b, int firstItem, int nItems, int parity, hipStream_t st)
{
    if (nItems <= 0) return hipSuccess;
    const int count = b->numGrids;
    const BatchDev* g = b->dGrids;
#define CUSTEN_BATCH_CASE(TT, HH, VV, RR)                                              \
    if (b->H == HH && b->V == VV && b->rpw == RR)                                      \
        return launchOne<TT, HH, VV, RR>(b, count, g, firstItem, nItems, parity, st);
    if (b->dtype == Dtype::F64) {
        CUSTEN_BATCH_CASE(double, 5, 5, 8)
        CUSTEN_BATCH_CASE(double, 5, 5, 4)
        CUSTEN_BATCH_CASE(double, 3, 3, 8)
        CUSTEN_BATCH_CASE(double, 3, 3, 4)
    } else {
        CUSTEN_BATCH_CASE(float, 5, 5, 8)
        CUSTEN_BATCH_CASE(float, 5, 5, 4)
        CUSTEN_BATCH_CASE(float, 3, 3, 8)
        CUSTEN_BATCH_CASE(float, 3, 3, 4)
    }
#undef CUSTEN_BATCH_CASE
    return hipErrorInvalidValue;
}

// Strip tables for a tile configuration (host copy and device copy).
hipError_t configure(custenBatch* b, int rpw, int chunks)
{
    b->rpw = rpw;
    b->chunks = chunks;
    int64_t strips = 0;
    for (int g = 0; g < b->numGrids; g++) {
        b->host[g].stripBase = (int)strips;
        b->strips[g] = stripsFor(b->host[g].nx, b->host[g].ny, rpw, chunks);
        strips += b->strips[g];
    }
    if (strips > (int64_t)0x7fffffff) return hipErrorInvalidValue;
    b->totalStrips = (int)strips;
    return hipMemcpy(b->dGrids, b->host.data(), sizeof(BatchDev) * b->numGrids,
                     hipMemcpyHostToDevice);
}

// `steps` applies starting at parity p0, all grids in one launch per step.
hipError_t enqueueFused(const custenBatch* b, int steps, int p0, hipStream_t st)
{
    for (int s = 0; s < steps; s++) {
        hipError_t e = launchItems(b, 0, b->totalStrips, (p0 + s) & 1, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" custenBatch* custenBatchCreate(int deviceNum, int numGrids, const int* nx, const int* ny,
                                          void* const* dataOutput, void* const* dataInput,
                                          const double* weights, int numStenHoriz, int numStenLeft,
                                          int numStenVert, int numStenTop, int dtype)
{
    const char* fn = "custenBatchCreate";
    if (numGrids < 1 || !nx || !ny || !dataOutput || !dataInput || !weights) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": bad argument");
        return nullptr;
    }
    const bool shapeOK = (numStenHoriz == 5 && numStenVert == 5) || (numStenHoriz == 3 && numStenVert == 3);
    if (!shapeOK || numStenLeft < 0 || numStenLeft >= numStenHoriz || numStenTop < 0 ||
        numStenTop >= numStenVert) {
        setError(CUSTEN_ERROR_NOT_SUPPORTED,
                 std::string(fn) + ": batched grids support 3x3 and 5x5 stencils");
        return nullptr;
    }
    if (dtype != 0 && dtype != 1) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": dtype must be 0 (FP64) or 1 (FP32)");
        return nullptr;
    }
    hipError_t e = hipSetDevice(deviceNum);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return nullptr;
    }
    custenBatch* b = new custenBatch();
    b->device = deviceNum;
    b->numGrids = numGrids;
    b->dtype = dtype == 0 ? Dtype::F64 : Dtype::F32;
    b->H = numStenHoriz;
    b->L = numStenLeft;
    b->V = numStenVert;
    b->T = numStenTop;
    const size_t es = dtype == 0 ? 8 : 4;
    int64_t strips = 0;
    for (int g = 0; g < numGrids; g++) {
        if (nx[g] < 1 || ny[g] < 1 || !dataInput[g] || !dataOutput[g] || dataInput[g] == dataOutput[g]) {
            setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": bad grid " + std::to_string(g));
            delete b;
            return nullptr;
        }
        BatchDev d{};
        d.in = dataInput[g];
        d.out = dataOutput[g];
        d.n = nx[g];
        d.nx = nx[g];
        d.ny = ny[g];
        d.stripBase = (int)strips;
        d.vec = (reinterpret_cast<uintptr_t>(d.in) % (2 * es) == 0 &&
                 reinterpret_cast<uintptr_t>(d.out) % (2 * es) == 0 && nx[g] % 2 == 0)
                    ? 1
                    : 0;
        b->host.push_back(d);
        const int s = stripsFor(nx[g], ny[g], b->rpw, b->chunks);
        b->strips.push_back(s);
        strips += s;
    }
    if (strips > (int64_t)0x7fffffff) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": too many strips");
        delete b;
        return nullptr;
    }
    b->totalStrips = (int)strips;
    if ((e = hipMalloc(&b->dGrids, sizeof(BatchDev) * numGrids)) != hipSuccess ||
        (e = hipMemcpy(b->dGrids, b->host.data(), sizeof(BatchDev) * numGrids,
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMalloc(&b->dWeights, sizeof(double) * numStenHoriz * numStenVert)) != hipSuccess ||
        (e = hipMemcpy(b->dWeights, weights, sizeof(double) * numStenHoriz * numStenVert,
                       hipMemcpyDefault)) != hipSuccess ||
        (e = hipStreamCreate(&b->own)) != hipSuccess) {
        setHipError(e, fn);
        custenBatchDestroy(b);
        return nullptr;
    }
    return b;
}

extern "C" int custenBatchRun(custenBatch* b, int steps, int mode, hipStream_t stream)
{
    const char* fn = "custenBatchRun";
    if (!b || steps < 0 || (mode != 0 && mode != 1)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": bad argument");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    (void)hipSetDevice(b->device);
    hipStream_t st = stream ? stream : b->own;
    hipError_t e = hipSuccess;
    if (mode == 1) {
        // plain per-grid launch loop: one launch per grid per step
        for (int s = 0; s < steps && e == hipSuccess; s++) {
            const int par = (b->parity + s) & 1;
            for (int g = 0; g < b->numGrids && e == hipSuccess; g++)
                e = launchItems(b, b->host[g].stripBase, b->strips[g], par, st);
        }
    } else if (steps > 0) {
        if (!b->tuned && tunerEnabled()) {
            // trial applies write the output side without advancing the
            // state (the next real step overwrites it with the same values)
            TuneKey k;
            k.kind = 1;
            k.dtype = static_cast<int>(b->dtype);
            k.H = b->H;
            k.V = b->V;
            int64_t pts = 0;
            int maxSide = 0;
            for (const BatchDev& d : b->host) {
                pts += (int64_t)d.nx * d.ny;
                maxSide = std::max(maxSide, std::max(d.nx, d.ny));
            }
            k.nx = (int)(pts >> 20);
            k.ny = maxSide;
            k.count = b->numGrids;
            TuneVal v;
            if (!tunerLookup(k, v)) {
                static const std::vector<TuneVal> cands = {{4, 1}, {4, 4}, {8, 1}, {8, 2},
                                                           {8, 4}, {8, 8}};
                auto run = [&](const TuneVal& c) {
                    hipError_t err = hipSuccess;
                    if (c.rpw != b->rpw || c.chunks != b->chunks) {
                        err = hipStreamSynchronize(b->own);
                        if (err == hipSuccess) err = configure(b, c.rpw, c.chunks);
                    }
                    if (err != hipSuccess) return err;
                    return launchItems(b, 0, b->totalStrips, b->parity, b->own);
                };
                e = tunerMeasure(b->own, run, cands, 2, v);
                if (e == hipSuccess) tunerStore(k, v);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(b->own);
            if (e == hipSuccess) e = configure(b, v.rpw, v.chunks);
            if (e != hipSuccess) {
                setHipError(e, fn);
                return lastError();
            }
            b->tuned = true;
        }
        const auto key = std::make_pair(steps, b->parity);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            e = hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e2 = enqueueFused(b, steps, b->parity, b->own);
                e = hipStreamEndCapture(b->own, &graph);
                if (e2 != hipSuccess) e = e2;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                setHipError(e, fn);
                return lastError();
            }
            b->graphs[key] = exec;
            b->graphDefs[key] = graph;
            it = b->graphs.find(key);
        }
        e = hipGraphLaunch(it->second, st);
    }
    if (e != hipSuccess) {
        setHipError(e, fn);
        return lastError();
    }
    b->parity = (b->parity + steps) & 1;
    return CUSTEN_SUCCESS;
}

extern "C" int custenBatchParity(const custenBatch* b) { return b ? b->parity : -1; }

extern "C" void custenBatchDestroy(custenBatch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->own) (void)hipStreamSynchronize(b->own);
    (void)hipDeviceSynchronize();
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : b->graphDefs) (void)hipGraphDestroy(kv.second);
    if (b->dGrids) (void)hipFree(b->dGrids);
    if (b->dWeights) (void)hipFree(b->dWeights);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
}
