// plan_internal.h -- library-private state behind cuSten_t::plan and the
// error/last-error machinery shared by all host entry points.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/cuSten.h"
#include "../../include/custen_point.h"
#include "kernels.h"

namespace custen {

enum class Dir : int { X = 0, Y = 1, XY = 2 };

// Where a data pointer lives, as far as a kernel launch is concerned.
enum class MemKind : int { Device = 0, Host = 1 };

struct Plan {
    Dir dir = Dir::XY;
    bool periodic = true;
    bool fun = false;
    Dtype dtype = Dtype::F64;
    int H = 1, L = 0, V = 1, T = 0;
    int numCoe = 0;
    Builtin builtin = Builtin::None;
    bool forceIndirect = false;
    custenPointLauncher pointLauncher = nullptr;  // caller-compiled point kernel
    bool zeroRight = false;     // x-np: right R columns written as 0.0
    bool weno = false;          // WENO5 advection (cuStenCreate2DXYWENOADVp)
    bool wenoSingleSquares = true;  // smoothness squares in single precision (as powf)
    hipStream_t userStream = nullptr;
    bool hasUserStream = false;
    int rowsPerWave = 0;        // set by custenSetTileConfig[Ex] (0: autotuned)
    int chunks = 0;
    int colsPerLane = 0;
    bool band = false;          // set by custenSetBandConfig
    bool tuned = false;         // autotuner result for resident launches
    int tunedRpw = 0, tunedChunks = 0, tunedCpl = 0;
    bool tunedBand = false;
    // device copy of host-resident weights / coefficients
    double* paramDev = nullptr;
    size_t paramCap = 0;
    // streaming resources (host-resident data)
    void* stageIn[2] = {nullptr, nullptr};
    void* stageOut[2] = {nullptr, nullptr};
    void* stageSide[2] = {nullptr, nullptr};  // WENO: u and v rows of a tile
    size_t stageInBytes = 0, stageOutBytes = 0, stageSideBytes = 0;
    hipEvent_t evH2D[2] = {nullptr, nullptr};
    hipEvent_t evComp[2] = {nullptr, nullptr};
    hipEvent_t evD2H[2] = {nullptr, nullptr};
    bool streamEventsReady = false;
    // pointer -> memory kind, so that repeated Computes on the same buffers
    // skip the runtime's pointer query
    const void* kindPtr[8] = {};
    MemKind kind[8] = {};
    int kindNext = 0;
    uint32_t magic = 0x43535450u;  // "CSTP"
};

// ---- errors ----------------------------------------------------------------
void setError(int code, const std::string& msg);
void setHipError(hipError_t e, const char* what);
int lastError();

// Classify a pointer; clears any HIP error the query leaves behind.
MemKind classify(const void* p, int device);
// The same through the plan's small cache of recently classified pointers.
MemKind classifyCached(Plan* plan, const void* p, int device);

// Validated access to a handle's plan (records an error and returns null).
Plan* planOf(cuSten_t* h, const char* fn);

}  // namespace custen
