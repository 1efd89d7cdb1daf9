// ch_noalloc.hip -- custenCHAdvanceNoAlloc: the fused Cahn-Hilliard ADI step
// of custenCHRun (mode 0) for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).  The caller passes the
// two time levels; everything else the step kernels use -- the descriptor,
// the factored diagonals and the two n x n work buffers -- lives in one
// file-scope device array, `scratch`, of fixed size (n <= kNoAllocMaxN).
//
// The factors are computed on the device by one thread from the same source
// as the host setup (ch_setup.h), so the results are bit-identical to
// custenCHRun's, and the call issues no host-to-device copy: the descriptor's
// constants travel as a kernel argument.  The call allocates nothing: no
// hipMalloc, hipFree or allocating library routine.
//
// Every call on one GPU shares that GPU's array and overwrites it; calls that
// overlap on one GPU (two host threads, or two streams that run at once)
// race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/cuSten.h"
#include "../../include/custen_ch.h"
#include "ch_kernels.h"
#include "ch_setup.h"
#include "plan_internal.h"

#pragma clang fp contract(off)

namespace custen {
namespace {

constexpr int kNoAllocMaxN = 1024;
// layout in doubles: descriptor | factors (7 m) | X (n*n) | Y (n*n)
constexpr size_t kDescDoubles = 128;
constexpr size_t kFacDoubles = ((size_t)7 * (kNoAllocMaxN - 2) + 31) / 32 * 32;
constexpr size_t kFieldDoubles = (size_t)kNoAllocMaxN * kNoAllocMaxN;
constexpr size_t kScratchDoubles = kDescDoubles + kFacDoubles + 2 * kFieldDoubles;
static_assert(sizeof(CHDesc) <= kDescDoubles * sizeof(double), "descriptor slot too small");

}  // namespace

// The work array, one copy per GPU, held until the process ends.  In host
// code `scratch` names a host-side stand-in: its device address on the
// current GPU comes from hipGetSymbolAddress (scratchBase).
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current; looked up
// once per device.
hipError_t scratchBase(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        g_scratchPtr[device] = static_cast<double*>(p);
    }
    *out = g_scratchPtr[device];
    return hipSuccess;
}

// Factor the matrix into fac (7 m doubles), form omega, and write the
// completed descriptor to *out.  One thread runs the recurrences on a copy in
// LDS (7 m doubles of dynamic shared memory); the workgroup then writes the
// factors out.
__global__ __launch_bounds__(256) void setupKernel(CHDesc d, double cDiag, double* fac,
                                                   CHDesc* out)
{
    extern __shared__ double lf[];
    const int m = d.n - 2;
    if (threadIdx.x == 0) {
        chFactor(lf, m, d.a, d.b, cDiag, d.d, d.e, d.omega);
        int cf = 0, cb = 0;
        chConstantRuns(lf, m, cf, cb);
        d.constFwd1 = cf + 1;
        d.constBwd1 = cb + 1;
        d.fastDiv = chDivCertified(lf[2 * m + cf]) ? 1 : 0;
        *out = d;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 7 * m; i += blockDim.x) fac[i] = lf[i];
}

// Exchange the contents of a and b (count elements each).
__global__ __launch_bounds__(256) void swapKernel(double* __restrict__ a, double* __restrict__ b,
                                                  int64_t count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double x = a[i];
        a[i] = b[i];
        b[i] = x;
    }
}

}  // namespace
}  // namespace custen

using namespace custen;

extern "C" int custenCHAdvanceNoAlloc(int deviceNum, const custenCHParams* params, double* c,
                                      double* cOld, int steps, hipStream_t stream)
{
    static const char* fn = "custenCHAdvanceNoAlloc";
    auto fail = [&](int code, const std::string& why) {
        setError(code, std::string(fn) + ": " + why);
        return code;
    };
    if (params == nullptr) return fail(CUSTEN_ERROR_INVALID_VALUE, "null params");
    if (c == nullptr || cOld == nullptr) return fail(CUSTEN_ERROR_INVALID_VALUE, "null field");
    if (c == cOld) return fail(CUSTEN_ERROR_INVALID_VALUE, "c and cOld must not alias");
    if (steps < 0) return fail(CUSTEN_ERROR_INVALID_VALUE, "steps must be >= 0");
    const int n = params->n;
    if (n < 8 || n > kNoAllocMaxN)
        return fail(CUSTEN_ERROR_INVALID_VALUE,
                    "n = " + std::to_string(n) + " outside [8, " + std::to_string(kNoAllocMaxN) +
                        "], the size of the allocation-free work array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || deviceNum < 0 || deviceNum >= ndev ||
        deviceNum >= kMaxDevices) {
        (void)hipGetLastError();
        return fail(CUSTEN_ERROR_INVALID_VALUE, "invalid deviceNum");
    }
    CHDesc d;
    double cDiag = 0.0;
    std::string why;
    if (!chDescribe(*params, d, cDiag, why)) return fail(CUSTEN_ERROR_INVALID_VALUE, why);
    hipError_t e = hipSetDevice(deviceNum);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return CUSTEN_ERROR_HIP;
    }
    // both time levels must be device memory of this GPU (no staging here)
    for (const double* p : {static_cast<const double*>(c), static_cast<const double*>(cOld)})
        if (classify(p, deviceNum) != MemKind::Device)
            return fail(CUSTEN_ERROR_INVALID_VALUE, "c and cOld must be device memory");
    double* base = nullptr;
    if ((e = scratchBase(deviceNum, &base)) != hipSuccess) {
        setHipError(e, fn);
        return CUSTEN_ERROR_HIP;
    }
    if (steps == 0) return CUSTEN_SUCCESS;

    CHDesc* dDesc = reinterpret_cast<CHDesc*>(base);
    double* fac = base + kDescDoubles;
    // two field buffers: each step writes the new field over the previous
    // one (the final phase is pointwise), so no third n*n array is needed
    d.f[0] = c;
    d.f[1] = cOld;
    d.f[2] = nullptr;
    d.ring = 2;
    d.X = fac + kFacDoubles;
    d.Y = d.X + kFieldDoubles;
    d.fac = fac;
    d.admitStep = 0;
    d.tileBase = d.colBase = d.trBase = 0;
    const size_t lds = sizeof(double) * 7 * (size_t)(n - 2);  // <= 57232 bytes
    hipLaunchKernelGGL(setupKernel, dim3(1), dim3(256), lds, stream, d, cDiag, fac, dDesc);
    if ((e = hipGetLastError()) != hipSuccess) {
        setHipError(e, fn);
        return CUSTEN_ERROR_HIP;
    }
    const int tiles = chTiles(n), trTiles = chTrTiles(n), colBlocks = chColBlocks(n);
    for (int s = 0; s < steps; s++) {
        if ((e = chLaunchRhs(dDesc, 1, tiles, s, stream)) != hipSuccess ||
            (e = chLaunchColSolve(dDesc, 1, colBlocks, 0, stream)) != hipSuccess ||
            (e = chLaunchTranspose(dDesc, 1, trTiles, stream)) != hipSuccess ||
            (e = chLaunchColSolve(dDesc, 1, colBlocks, 1, stream)) != hipSuccess ||
            (e = chLaunchFinal(dDesc, 1, tiles, s, stream)) != hipSuccess) {
            setHipError(e, fn);
            return CUSTEN_ERROR_HIP;
        }
    }
    // step s writes the new field over f[(s + 1) & 1]: after an odd number of
    // steps the current field is in cOld's buffer
    if (steps & 1) {
        const int64_t nn = (int64_t)n * n;
        const int blocks = (int)((nn + 255) / 256 < 4096 ? (nn + 255) / 256 : 4096);
        hipLaunchKernelGGL(swapKernel, dim3(blocks), dim3(256), 0, stream, c, cOld, nn);
        if ((e = hipGetLastError()) != hipSuccess) {
            setHipError(e, fn);
            return CUSTEN_ERROR_HIP;
        }
    }
    return CUSTEN_SUCCESS;
}
