// batch.hip -- many independent periodic xy grids of uneven sizes advanced
// together: one launch per step covers every strip of every grid (the
// workgroup finds its grid by binary search over the strip prefix sums), and
// a whole multi-step run is captured once into a HIP graph and replayed.
// With the band walker selected, a step is two launches: the band walker over
// the interior 128-column strips of every grid (one wave per band, found the
// same way), then the strip kernel over the remaining edge strips.  The
// per-grid launch loop (one launch per grid per step) is kept as the baseline
// mode.  Per grid the arithmetic is the single-grid kernels' (stencil_body.h,
// stencil_band.h), so results are bit-identical to cuStenCompute2DXYp.
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
#include "stencil_band.h"
#include "stencil_body.h"
#include "tuner.h"

#pragma clang fp contract(off)

namespace custen {

struct BatchDev {
    const void* in;
    void* out;
    int n;          // nx == ny == n for square grids; nx, ny kept separately below
    int nx, ny;
    int stripBase;  // first strip id of this grid (strip kernel over all strips)
    int vec;        // rows pair-aligned
    // band-walker split: strips [0, nBandTx) are walked by bands, strips
    // [nBandTx, tilesX) run the strip kernel
    int bandBase;   // first band item (wave) of this grid
    int nBandTx;
    int edgeBase;   // first edge-strip item of this grid
};

namespace dev {

// The single-grid launch description of grid G (periodic xy, all points
// written), reading the side given by the step parity.
__device__ __forceinline__ StencilLaunch gridLaunch(const BatchDev& G, int parity, Dtype dt,
                                                    const double* weights, int H, int L, int V,
                                                    int Tt, int rpw, int chunks)
{
    StencilLaunch a;
    a.dtype = dt;
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
    a.rowsPerWave = rpw;
    a.chunks = chunks;
    a.vecIn = G.vec;
    a.vecOut = G.vec;
    a.band = 0;
    a.stripTx0 = 0;
    a.nStripTx = 0;
    a.edgeList = 0;
    return a;
}

// Strip kernel over every strip of every grid (edges = 0) or over the edge
// strips [nBandTx, tilesX) of every grid (edges = 1).
// itemGrid[item] is the grid that owns the item (built on the host with the
// item tables, so a workgroup finds its grid with one load).
template <typename T, int H, int V, int RPW>
__global__ __launch_bounds__(kThreads) void batchKernel(const BatchDev* __restrict__ grids,
                                                        const int* __restrict__ itemGrid,
                                                        int firstItem, int parity,
                                                        const double* __restrict__ weights, int L,
                                                        int Tt, int chunks, int edges)
{
    __shared__ __attribute__((aligned(16))) T tile[(kWaves * RPW + V - 1) * ldsRowLength(H)];
    T w[H * V];
#pragma unroll
    for (int k = 0; k < H * V; k++) w[k] = static_cast<T>(weights[k]);
    const int item = firstItem + xcdRemap(blockIdx.x, gridDim.x);
    const BatchDev& G = grids[itemGrid[item]];
    const StencilLaunch a = gridLaunch(G, parity, sizeof(T) == 8 ? Dtype::F64 : Dtype::F32,
                                       weights, H, L, V, Tt, RPW, chunks);
    const int tilesX = (G.nx + kTW - 1) / kTW;
    const int tx0 = edges ? G.nBandTx : 0;
    const int k = item - (edges ? G.edgeBase : G.stripBase);
    const int tx = tx0 + k % (tilesX - tx0), tyg = k / (tilesX - tx0);
    if (stripIsInterior(a, tx)) stencilStripImpl<T, H, V, RPW, Xform::Identity, true>(a, w, tile, tx, tyg);
    else stencilStripImpl<T, H, V, RPW, Xform::Identity, false>(a, w, tile, tx, tyg);
}

// Band walker over the interior strips [0, nBandTx) of every grid.
template <typename T, int H, int V, int RPW>
__global__ __launch_bounds__(kLanes) void batchBandKernel(const BatchDev* __restrict__ grids,
                                                          const int* __restrict__ itemGrid,
                                                          int parity,
                                                          const double* __restrict__ weights,
                                                          int L, int Tt, int chunks)
{
    using BS = BandShape<T, H, V, RPW>;
    __shared__ __attribute__((aligned(16))) T ring[BS::S * BS::LD];
    T w[H * V];
#pragma unroll
    for (int k = 0; k < H * V; k++) w[k] = static_cast<T>(weights[k]);
    const int item = xcdRemap(blockIdx.x, gridDim.x);
    const BatchDev& G = grids[itemGrid[item]];
    StencilLaunch a = gridLaunch(G, parity, sizeof(T) == 8 ? Dtype::F64 : Dtype::F32, weights, H,
                                 L, V, Tt, RPW, chunks);
    a.nStripTx = G.nBandTx;
    Lds<T>* r = (Lds<T>*)ring;
    if (G.ny % RPW == 0) bandWalk<T, H, V, RPW, true>(a, w, r, item - G.bandBase);
    else bandWalk<T, H, V, RPW, false>(a, w, r, item - G.bandBase);
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
    bool band = false;        // fused steps use the band walker for interior strips
    std::vector<BatchDev> host;
    std::vector<int> strips;  // strips per grid
    int totalStrips = 0;
    int totalBand = 0, totalEdge = 0;  // band items and edge-strip items (band mode)
    BatchDev* dGrids = nullptr;
    // owning grid of every item: all strips, band items, edge-strip items
    int* dItemGrid[3] = {nullptr, nullptr, nullptr};
    size_t itemCap[3] = {0, 0, 0};
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
hipError_t launchOne(const custenBatch* b, const BatchDev* grids, int firstItem, int nItems,
                     int parity, int kind, hipStream_t st)
{
    if (kind == 2)
        hipLaunchKernelGGL((dev::batchBandKernel<T, H, V, RPW>), dim3(nItems), dim3(dev::kLanes),
                           0, st, grids, b->dItemGrid[2], parity, b->dWeights, b->L, b->T,
                           b->chunks);
    else
        hipLaunchKernelGGL((dev::batchKernel<T, H, V, RPW>), dim3(nItems), dim3(dev::kThreads), 0,
                           st, grids, b->dItemGrid[kind], firstItem, parity, b->dWeights, b->L,
                           b->T, b->chunks, kind);
    return hipGetLastError();
}

// kind 0: strip kernel over items [firstItem, +nItems) of the all-strips
// table; 1: strip kernel over the edge-strip table; 2: band walker over the
// band table (firstItem must be 0 for kinds 1 and 2).
hipError_t launchItems(const custenBatch* b, int firstItem, int nItems, int parity, hipStream_t st,
                       int kind = 0)
{
    if (nItems <= 0) return hipSuccess;
    const BatchDev* g = b->dGrids;
#define CUSTEN_BATCH_CASE(TT, HH, VV, RR)                                              \
    if (b->H == HH && b->V == VV && b->rpw == RR)                                      \
        return launchOne<TT, HH, VV, RR>(b, g, firstItem, nItems, parity, kind, st);
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

// Item tables for a tile configuration (host copy and device copy): all
// strips (per-grid loop and strip-only fused steps) and, with the band
// walker, its band items and the remaining edge strips.  A grid's strips are
// walked by bands when its rows are pair-aligned and its height allows one
// wrap correction per row (ny >= V + rows per wave).
hipError_t configure(custenBatch* b, int rpw, int chunks, bool band)
{
    b->rpw = rpw;
    b->chunks = chunks;
    b->band = band;
    int64_t strips = 0, bands = 0, edges = 0;
    for (int g = 0; g < b->numGrids; g++) {
        BatchDev& d = b->host[g];
        d.stripBase = (int)strips;
        b->strips[g] = stripsFor(d.nx, d.ny, rpw, chunks);
        strips += b->strips[g];
        const int tilesX = (d.nx + dev::kTW - 1) / dev::kTW;
        d.nBandTx = (band && d.vec && d.ny >= b->V + rpw) ? d.nx / dev::kTW : 0;
        d.bandBase = (int)bands;
        d.edgeBase = (int)edges;
        const int bandRows = rpw * chunks;
        const int groupRows = dev::kWaves * rpw * chunks;
        bands += (int64_t)d.nBandTx * ((d.ny + bandRows - 1) / bandRows);
        edges += (int64_t)(tilesX - d.nBandTx) * ((d.ny + groupRows - 1) / groupRows);
    }
    if (strips > (int64_t)0x7fffffff || bands > (int64_t)0x7fffffff ||
        edges > (int64_t)0x7fffffff)
        return hipErrorInvalidValue;
    b->totalStrips = (int)strips;
    b->totalBand = (int)bands;
    b->totalEdge = (int)edges;
    // owning grid of every item of the three tables
    std::vector<int> owner[3];
    owner[0].reserve(b->totalStrips);
    owner[1].reserve(b->totalEdge);
    owner[2].reserve(b->totalBand);
    for (int g = 0; g < b->numGrids; g++) {
        const BatchDev& d = b->host[g];
        const int nBand = (g + 1 < b->numGrids ? b->host[g + 1].bandBase : b->totalBand) - d.bandBase;
        const int nEdge = (g + 1 < b->numGrids ? b->host[g + 1].edgeBase : b->totalEdge) - d.edgeBase;
        owner[0].insert(owner[0].end(), b->strips[g], g);
        owner[1].insert(owner[1].end(), nEdge, g);
        owner[2].insert(owner[2].end(), nBand, g);
    }
    for (int k = 0; k < 3; k++) {
        const size_t n = std::max<size_t>(1, owner[k].size());
        if (n > b->itemCap[k]) {
            if (b->dItemGrid[k]) (void)hipFree(b->dItemGrid[k]);
            b->dItemGrid[k] = nullptr;
            b->itemCap[k] = 0;
            hipError_t e = hipMalloc(&b->dItemGrid[k], n * sizeof(int));
            if (e != hipSuccess) return e;
            b->itemCap[k] = n;
        }
        if (!owner[k].empty()) {
            hipError_t e = hipMemcpy(b->dItemGrid[k], owner[k].data(), owner[k].size() * sizeof(int),
                                     hipMemcpyHostToDevice);
            if (e != hipSuccess) return e;
        }
    }
    return hipMemcpy(b->dGrids, b->host.data(), sizeof(BatchDev) * b->numGrids,
                     hipMemcpyHostToDevice);
}

// One fused step at parity p: one launch over all strips, or the band walker
// then the edge strips.
hipError_t enqueueStep(const custenBatch* b, int p, hipStream_t st)
{
    if (!b->band) return launchItems(b, 0, b->totalStrips, p, st, 0);
    hipError_t e = launchItems(b, 0, b->totalBand, p, st, 2);
    if (e == hipSuccess) e = launchItems(b, 0, b->totalEdge, p, st, 1);
    return e;
}

// `steps` applies starting at parity p0.
hipError_t enqueueFused(const custenBatch* b, int steps, int p0, hipStream_t st)
{
    for (int s = 0; s < steps; s++) {
        hipError_t e = enqueueStep(b, (p0 + s) & 1, st);
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
        (e = configure(b, b->rpw, b->chunks, false)) != hipSuccess ||
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
                // strip kernel (rows per wave, chunks per workgroup), then
                // band walker (rows per wave, chunks per band)
                static const std::vector<TuneVal> cands = {
                    {4, 1}, {4, 4}, {8, 1}, {8, 2}, {8, 4}, {8, 8},
                    {4, 2, 2, 1}, {4, 4, 2, 1}, {8, 1, 2, 1}, {8, 2, 2, 1}};
                auto run = [&](const TuneVal& c) {
                    hipError_t err = hipSuccess;
                    if (c.rpw != b->rpw || c.chunks != b->chunks || (c.band != 0) != b->band) {
                        err = hipStreamSynchronize(b->own);
                        if (err == hipSuccess) err = configure(b, c.rpw, c.chunks, c.band != 0);
                    }
                    if (err != hipSuccess) return err;
                    return enqueueStep(b, b->parity, b->own);
                };
                e = tunerMeasure(b->own, run, cands, 2, v);
                if (e == hipSuccess) tunerStore(k, v);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(b->own);
            if (e == hipSuccess) e = configure(b, v.rpw, v.chunks, v.band != 0);
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

extern "C" int custenBatchSetTileConfig(custenBatch* b, int rowsPerWave, int chunks, int band)
{
    const char* fn = "custenBatchSetTileConfig";
    if (!b || (rowsPerWave != 4 && rowsPerWave != 8) || chunks < 1 || (band != 0 && band != 1)) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 std::string(fn) + ": rowsPerWave must be 4 or 8, chunks >= 1, band 0 or 1");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    (void)hipSetDevice(b->device);
    // the device tables are rewritten: no earlier step may still read them,
    // and graphs captured for the old item counts are discarded
    hipError_t e = hipStreamSynchronize(b->own);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto& kv : b->graphDefs) (void)hipGraphDestroy(kv.second);
        b->graphs.clear();
        b->graphDefs.clear();
        e = configure(b, rowsPerWave, chunks, band != 0);
    }
    if (e != hipSuccess) {
        setHipError(e, fn);
        return lastError();
    }
    b->tuned = true;
    return CUSTEN_SUCCESS;
}

extern "C" int custenBatchGetTileConfig(const custenBatch* b, int* rowsPerWave, int* chunks,
                                        int* band)
{
    if (!b) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenBatchGetTileConfig: null batch");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    if (rowsPerWave) *rowsPerWave = b->tuned ? b->rpw : 0;
    if (chunks) *chunks = b->tuned ? b->chunks : 0;
    if (band) *band = b->tuned && b->band ? 1 : 0;
    return CUSTEN_SUCCESS;
}

extern "C" void custenBatchDestroy(custenBatch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->own) (void)hipStreamSynchronize(b->own);
    (void)hipDeviceSynchronize();
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : b->graphDefs) (void)hipGraphDestroy(kv.second);
    if (b->dGrids) (void)hipFree(b->dGrids);
    for (int k = 0; k < 3; k++)
        if (b->dItemGrid[k]) (void)hipFree(b->dItemGrid[k]);
    if (b->dWeights) (void)hipFree(b->dWeights);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
}
