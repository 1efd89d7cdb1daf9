// fun_caller.cpp -- a C++ program with its OWN device point functions, shaped
// like the reference's Fun examples and its Cahn-Hilliard program
// (examples/src/2d_xy_p_fun.cu:56-80,160-200, examples/src/2d_x_p_fun.cu,
// cuPentCahnADI/src/cuPentCahnADI.cu:164-190,486-516):
//
//     __device__ devArg1XY devFunc = f;                      // this file
//     hipMemcpyFromSymbol(&func, HIP_SYMBOL(devFunc), ...);  // fetch the pointer
//     cuStenCreate2DXYpFun(..., coe, func);                  // library calls it
//
// It is compiled with -fgpu-rdc and linked against libcusten.a, so the
// library's kernels and these functions share one GPU code object (the
// counterpart of the reference's nvcc --device-c, cuSten/Makefile:20).  The
// GPU tests run it as a child process and compare what it writes with the
// oracle.  The sums use explicit fma, as nvcc contracts `result += c * d`
// in the reference's functions; everything else is left unfused
// (-ffp-contract=off on this file).
//
// usage: fun_caller <nx> <ny> <seed> <out.bin>
//   writes three nx*ny FP64 fields, in this order:
//     1. x periodic,  5-point weighted sum  (devArg1X,  xWeightedSum5)
//     2. xy periodic, 3x3 weighted sum      (devArg1XY, centralDiff as in 2d_xy_p_fun.cu)
//     3. xy non-periodic, 3x3 sum of coe * (u^3 - u)  (devArg1XY, nonLinRHS)
// usage: fun_caller time <n> <reps>
//   times `reps` xy-periodic applies of each of this file's xy functions on
//   an n x n device-resident field (HIP events, after warm-up applies) and
//   prints one JSON line with the microseconds per apply
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include <hip/hip_runtime.h>

#include "cuSten.h"

// devArg1X: loc is the window centre; five points, two on each side
__device__ double xWeightedSum5(double* data, double* coe, int loc)
{
    double result = 0.0;
    for (int k = 0; k < 5; k++) result = fma(coe[k], data[loc - 2 + k], result);
    return result;
}

// devArg1XY: loc is the window's top-left point, jump the row stride
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

}  // namespace

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
    double* funcs[2] = {nullptr, nullptr};
    check(hipMemcpyFromSymbol(&funcs[0], HIP_SYMBOL(devFuncNonLin), sizeof(devArg1XY)), "fetch");
    check(hipMemcpyFromSymbol(&funcs[1], HIP_SYMBOL(devFuncXY), sizeof(devArg1XY)), "fetch");
    const char* names[2] = {"nonLinRHS", "centralDiff"};
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    std::printf("{\"n\": %d, \"reps\": %d", n, reps);
    for (int k = 0; k < 2; k++) {
        cuSten_t h;
        cuStenCreate2DXYpFun(&h, 0, 1, n, n, 32, 32, out, in, coe, 3, 1, 1, 3, 1, 1, funcs[k]);
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

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "time")
        return timeApplies(std::atoi(argv[2]), std::atoi(argv[3]));
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s nx ny seed out.bin\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]);
    const int ny = std::atoi(argv[2]);
    const unsigned long long seed = std::strtoull(argv[3], nullptr, 10);
    const int deviceNum = 0, numTiles = 1;
    const size_t n = (size_t)nx * ny;

    double *dataInput = nullptr, *dataOutput = nullptr, *coe = nullptr;
    check(hipMallocManaged(&dataInput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&dataOutput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&coe, 9 * sizeof(double)), "hipMallocManaged");
    custenFillUniform(dataInput, n, seed, -1.0, 1.0);
    custenFillUniform(coe, 9, seed + 1, -1.0, 1.0);

    // the device function pointers, fetched as the reference's callers do
    double* funcX = nullptr;
    double* funcXY = nullptr;
    double* funcNonLin = nullptr;
    check(hipMemcpyFromSymbol(&funcX, HIP_SYMBOL(devFuncX), sizeof(devArg1X)), "fetch devFuncX");
    check(hipMemcpyFromSymbol(&funcXY, HIP_SYMBOL(devFuncXY), sizeof(devArg1XY)),
          "fetch devFuncXY");
    check(hipMemcpyFromSymbol(&funcNonLin, HIP_SYMBOL(devFuncNonLin), sizeof(devArg1XY)),
          "fetch devFuncNonLin");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");

    FILE* f = std::fopen(argv[4], "wb");
    if (!f) {
        std::perror("fopen");
        return 2;
    }

    // 1. x periodic: cuStenCreate2DXpFun(..., coe, numCoe, func)
    for (size_t i = 0; i < n; i++) dataOutput[i] = 0.0;
    cuSten_t xDirCompute;
    cuStenCreate2DXpFun(&xDirCompute, deviceNum, numTiles, nx, ny, 32, 16, dataOutput, dataInput,
                        coe, 5, 2, 2, 5, funcX);
    checkError("create x");
    cuStenCompute2DXpFun(&xDirCompute, HOST);
    checkError("compute x");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXpFun(&xDirCompute);

    // 2. xy periodic weighted sum (examples/src/2d_xy_p_fun.cu)
    for (size_t i = 0; i < n; i++) dataOutput[i] = 0.0;
    cuSten_t xyDirCompute;
    cuStenCreate2DXYpFun(&xyDirCompute, deviceNum, numTiles, nx, ny, 16, 32, dataOutput,
                         dataInput, coe, 3, 1, 1, 3, 1, 1, funcXY);
    checkError("create xy");
    cuStenCompute2DXYpFun(&xyDirCompute, HOST);
    checkError("compute xy");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXYpFun(&xyDirCompute);

    // 3. xy non-periodic nonlinear term (cuPentCahnADI.cu:486-516 shape)
    for (size_t i = 0; i < n; i++) dataOutput[i] = 0.0;
    cuSten_t nonLinCompute;
    cuStenCreate2DXYnpFun(&nonLinCompute, deviceNum, numTiles, nx, ny, 16, 32, dataOutput,
                          dataInput, coe, 3, 1, 1, 3, 1, 1, funcNonLin);
    checkError("create nonlin");
    cuStenCompute2DXYnpFun(&nonLinCompute, HOST);
    checkError("compute nonlin");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::fwrite(dataOutput, sizeof(double), n, f);
    cuStenDestroy2DXYnpFun(&nonLinCompute);

    std::fclose(f);
    (void)hipFree(dataInput);
    (void)hipFree(dataOutput);
    (void)hipFree(coe);
    std::printf("fun_caller: %dx%d done\n", nx, ny);
    return 0;
}
