// point_caller.cpp -- a C++ program whose point functions are compiled into
// its own stencil kernels (include/custen_point.h) instead of being called
// through a device function pointer.  It creates the reference's Fun handles
// exactly as examples/src/2d_xy_p_fun.cu does (the function pointer fetched
// with hipMemcpyFromSymbol) and then attaches the launcher instantiated for
// the same function.  Compiled without relocatable device code and linked
// against libcusten.so.  The GPU tests run it as a child process and compare
// what it writes with the oracle.  The sums use explicit fma as nvcc
// contracts the reference's `result += c * d`; everything else is unfused
// (-ffp-contract=off).
//
// usage: point_caller <nx> <ny> <seed> <out.bin>
//   writes five nx*ny FP64 fields, in this order:
//     1. x periodic,  5-point weighted sum (devArg1X)
//     2. y non-periodic, 5-point weighted sum (devArg1Y)
//     3. xy periodic, 3x3 weighted sum (devArg1XY)
//     4. xy non-periodic, 3x3 sum of coe * (u^3 - u) (devArg1XY)
//     5. as 3, with input and output in pinned host memory, numTiles = 4
// usage: point_caller time <n> <reps>
//   times `reps` xy-periodic applies of both xy functions on an n x n
//   device-resident field and prints one JSON line (microseconds per apply)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include <hip/hip_runtime.h>

#include "cuSten.h"
#include "custen_point.h"

__device__ double xWeightedSum5(double* data, double* coe, int loc)
{
    double result = 0.0;
    for (int k = 0; k < 5; k++) result = fma(coe[k], data[loc - 2 + k], result);
    return result;
}

__device__ double yWeightedSum5(double* data, double* coe, int loc, int jump)
{
    double result = 0.0;
    for (int k = 0; k < 5; k++) result = fma(coe[k], data[loc + (k - 2) * jump], result);
    return result;
}

__device__ double centralDiff(double* data, double* coe, int loc, int jump, int nx, int ny)
{
    double result = 0.0;
    int count = 0;
    for (int j = 0; j < ny; j++) {
        const int temp = loc + j * jump;
        for (int i = 0; i < nx; i++) {
            result = fma(coe[count], data[temp + i], result);
            count++;
        }
    }
    return result;
}

__device__ double nonLinRHS(double* data, double* coe, int loc, int jump, int nx, int ny)
{
    double result = 0.0;
    int count = 0;
    for (int j = 0; j < ny; j++) {
        const int temp = loc + j * jump;
        for (int i = 0; i < nx; i++) {
            const double current = data[temp + i];
            result = fma(coe[count], (current * current * current) - current, result);
            count++;
        }
    }
    return result;
}

__device__ devArg1X devFuncX = xWeightedSum5;
__device__ devArg1Y devFuncY = yWeightedSum5;
__device__ devArg1XY devFuncXY = centralDiff;
__device__ devArg1XY devFuncNonLin = nonLinRHS;

namespace {

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

// the device function pointer held by a __device__ variable of this file
template <class Symbol>
double* fetch(const Symbol& symbol)
{
    double* f = nullptr;
    check(hipMemcpyFromSymbol(&f, symbol, sizeof(double*)), "hipMemcpyFromSymbol");
    return f;
}

int timeApplies(int n, int reps)
{
    const size_t pts = (size_t)n * n;
    double *in = nullptr, *out = nullptr, *coe = nullptr;
    check(hipMalloc(&in, pts * sizeof(double)), "hipMalloc");
    check(hipMalloc(&out, pts * sizeof(double)), "hipMalloc");
    check(hipMalloc(&coe, 9 * sizeof(double)), "hipMalloc");
    double* host = static_cast<double*>(std::malloc(pts * sizeof(double)));
    custenFillUniform(host, pts, 2, -1.0, 1.0);
    check(hipMemcpy(in, host, pts * sizeof(double), hipMemcpyHostToDevice), "upload");
    std::free(host);
    double hc[9];
    custenFillUniform(hc, 9, 3, -1.0, 1.0);
    check(hipMemcpy(coe, hc, sizeof(hc), hipMemcpyHostToDevice), "upload");
    double* funcs[2] = {fetch(HIP_SYMBOL(devFuncNonLin)), fetch(HIP_SYMBOL(devFuncXY))};
    custenPointLauncher launchers[2] = {custenPointLauncherXY<nonLinRHS>,
                                        custenPointLauncherXY<centralDiff>};
    const char* names[2] = {"nonLinRHS", "centralDiff"};
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    std::printf("{\"n\": %d, \"reps\": %d", n, reps);
    for (int k = 0; k < 2; k++) {
        cuSten_t h;
        cuStenCreate2DXYpFun(&h, 0, 1, n, n, 32, 32, out, in, coe, 3, 1, 1, 3, 1, 1, funcs[k]);
        custenSetPointLauncher(&h, launchers[k]);
        checkError("create");
        for (int i = 0; i < 5; i++) cuStenCompute2DXYpFun(&h, DEVICE);
        check(hipDeviceSynchronize(), "sync");
        check(hipEventRecord(e0, h.streams[0]), "record");
        for (int i = 0; i < reps; i++) cuStenCompute2DXYpFun(&h, DEVICE);
        check(hipEventRecord(e1, h.streams[0]), "record");
        check(hipEventSynchronize(e1), "sync");
        checkError("compute");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        std::printf(", \"%s_us\": %.1f", names[k], 1e3 * ms / reps);
        cuStenDestroy2DXYpFun(&h);
    }
    std::printf("}\n");
    (void)hipFree(in);
    (void)hipFree(out);
    (void)hipFree(coe);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "time")
        return timeApplies(std::atoi(argv[2]), std::atoi(argv[3]));
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s nx ny seed out.bin | time n reps\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]);
    const int ny = std::atoi(argv[2]);
    const unsigned long long seed = std::strtoull(argv[3], nullptr, 10);
    const size_t n = (size_t)nx * ny;

    double *dataInput = nullptr, *dataOutput = nullptr, *coe = nullptr;
    check(hipMallocManaged(&dataInput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&dataOutput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&coe, 9 * sizeof(double)), "hipMallocManaged");
    custenFillUniform(dataInput, n, seed, -1.0, 1.0);
    custenFillUniform(coe, 9, seed + 1, -1.0, 1.0);
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");

    FILE* f = std::fopen(argv[4], "wb");
    if (!f) {
        std::perror("fopen");
        return 2;
    }
    auto clearOut = [&]() {
        for (size_t i = 0; i < n; i++) dataOutput[i] = 0.0;
    };

    // 1. x periodic
    clearOut();
    cuSten_t hx;
    cuStenCreate2DXpFun(&hx, 0, 1, nx, ny, 32, 16, dataOutput, dataInput, coe, 5, 2, 2, 5,
                        fetch(HIP_SYMBOL(devFuncX)));
    custenSetPointLauncher(&hx, custenPointLauncherX<xWeightedSum5>);
    cuStenCompute2DXpFun(&hx, HOST);
    checkError("x");
    check(hipDeviceSynchronize(), "sync");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXpFun(&hx);

    // 2. y non-periodic
    clearOut();
    cuSten_t hy;
    cuStenCreate2DYnpFun(&hy, 0, 1, nx, ny, 32, 16, dataOutput, dataInput, coe, 5, 2, 2,
                         fetch(HIP_SYMBOL(devFuncY)));
    custenSetPointLauncher(&hy, custenPointLauncherY<yWeightedSum5>);
    cuStenCompute2DYnpFun(&hy, HOST);
    checkError("y");
    check(hipDeviceSynchronize(), "sync");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DYnpFun(&hy);

    // 3. xy periodic weighted sum
    clearOut();
    cuSten_t hxy;
    cuStenCreate2DXYpFun(&hxy, 0, 1, nx, ny, 16, 32, dataOutput, dataInput, coe, 3, 1, 1, 3, 1, 1,
                         fetch(HIP_SYMBOL(devFuncXY)));
    custenSetPointLauncher(&hxy, custenPointLauncherXY<centralDiff>);
    cuStenCompute2DXYpFun(&hxy, HOST);
    checkError("xy");
    check(hipDeviceSynchronize(), "sync");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXYpFun(&hxy);

    // 4. xy non-periodic nonlinear term
    clearOut();
    cuSten_t hnl;
    cuStenCreate2DXYnpFun(&hnl, 0, 1, nx, ny, 16, 32, dataOutput, dataInput, coe, 3, 1, 1, 3, 1,
                          1, fetch(HIP_SYMBOL(devFuncNonLin)));
    custenSetPointLauncher(&hnl, custenPointLauncherXY<nonLinRHS>);
    cuStenCompute2DXYnpFun(&hnl, HOST);
    checkError("nonlin");
    check(hipDeviceSynchronize(), "sync");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXYnpFun(&hnl);

    // 5. xy periodic, host-resident pinned data streamed in 4 tiles
    double *hin = nullptr, *hout = nullptr;
    check(hipHostMalloc(&hin, n * sizeof(double), 0), "hipHostMalloc");
    check(hipHostMalloc(&hout, n * sizeof(double), 0), "hipHostMalloc");
    std::memcpy(hin, dataInput, n * sizeof(double));
    std::memset(hout, 0, n * sizeof(double));
    cuSten_t hs;
    cuStenCreate2DXYpFun(&hs, 0, 4, nx, ny, 16, 32, hout, hin, coe, 3, 1, 1, 3, 1, 1,
                         fetch(HIP_SYMBOL(devFuncXY)));
    custenSetPointLauncher(&hs, custenPointLauncherXY<centralDiff>);
    cuStenCompute2DXYpFun(&hs, HOST);
    checkError("streamed");
    check(hipDeviceSynchronize(), "sync");
    std::fwrite(hout, sizeof(double), n, f);
    cuStenDestroy2DXYpFun(&hs);
    (void)hipHostFree(hin);
    (void)hipHostFree(hout);

    std::fclose(f);
    (void)hipFree(dataInput);
    (void)hipFree(dataOutput);
    (void)hipFree(coe);
    std::printf("point_caller: %dx%d done\n", nx, ny);
    return 0;
}
