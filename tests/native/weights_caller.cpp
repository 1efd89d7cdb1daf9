// weights_caller.cpp -- a C++ program that uses the library the way the
// reference's example programs use cuSten (examples/src/2d_xy_p.cu:60-170):
// managed allocations filled on the host, cuStenCreate2DXYp, Compute,
// hipDeviceSynchronize, Swap, Compute again, Destroy, checkError after each
// call.  It links against libcusten.so and is run by the GPU tests as a child
// process, which compare what it writes with the oracle.
//
// usage: weights_caller <nx> <ny> <numTiles> <periodic 0|1> <seed> <out.bin>
//   input field  : SplitMix64 uniform [-1, 1) of `seed` (custenFillUniform)
//   weights      : 5 x 5, SplitMix64 uniform [-1, 1) of seed + 1
//   out.bin      : the field after two applies (apply, Swap, apply), nx*ny
//                  doubles, followed by the handle's pointer offsets after the
//                  Swap (see writeHandle)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cuSten.h"

namespace {

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

// Offsets (in elements, relative to `base`) of every per-tile pointer in the
// handle, so the test can check them against the reference's formulas.
void writeHandle(FILE* f, const cuSten_t& h, const double* base)
{
    std::vector<long long> v;
    for (int t = 0; t < h.numTiles; t++) {
        v.push_back(h.dataInput[t] - base);
        v.push_back(h.dataOutput[t] - base);
        v.push_back(h.boundaryTop ? h.boundaryTop[t] - base : -1);
        v.push_back(h.boundaryBottom ? h.boundaryBottom[t] - base : -1);
    }
    std::fwrite(v.data(), sizeof(long long), v.size(), f);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s nx ny numTiles periodic seed out.bin\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]);
    const int ny = std::atoi(argv[2]);
    const int numTiles = std::atoi(argv[3]);
    const int periodic = std::atoi(argv[4]);
    const unsigned long long seed = std::strtoull(argv[5], nullptr, 10);
    const int deviceNum = 0;

    // examples/src/2d_xy_p.cu:74-76: managed memory, initialised on the host
    double* dataInput = nullptr;
    double* dataOutput = nullptr;
    double* weights = nullptr;
    const size_t n = (size_t)nx * ny;
    check(hipMallocManaged(&dataInput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&dataOutput, n * sizeof(double)), "hipMallocManaged");
    check(hipMallocManaged(&weights, 25 * sizeof(double)), "hipMallocManaged");
    custenFillUniform(dataInput, n, seed, -1.0, 1.0);
    for (size_t i = 0; i < n; i++) dataOutput[i] = 0.0;
    custenFillUniform(weights, 25, seed + 1, -1.0, 1.0);
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");

    const int numStenHoriz = 5, numStenLeft = 2, numStenRight = 2;
    const int numStenVert = 5, numStenTop = 2, numStenBottom = 2;

    cuSten_t xyDirCompute;
    if (periodic) {
        cuStenCreate2DXYp(&xyDirCompute, deviceNum, numTiles, nx, ny, 32, 16, dataOutput,
                          dataInput, weights, numStenHoriz, numStenLeft, numStenRight, numStenVert,
                          numStenTop, numStenBottom);
    } else {
        cuStenCreate2DXYnp(&xyDirCompute, deviceNum, numTiles, nx, ny, 32, 16, dataOutput,
                           dataInput, weights, numStenHoriz, numStenLeft, numStenRight,
                           numStenVert, numStenTop, numStenBottom);
    }
    checkError("create");

    // two time steps: apply, swap the roles of the two fields, apply
    if (periodic) cuStenCompute2DXYp(&xyDirCompute, HOST);
    else cuStenCompute2DXYnp(&xyDirCompute, HOST);
    checkError("compute 1");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (periodic) cuStenSwap2DXYp(&xyDirCompute, dataOutput);
    else cuStenSwap2DXYnp(&xyDirCompute, dataOutput);
    checkError("swap");
    if (periodic) cuStenCompute2DXYp(&xyDirCompute, HOST);
    else cuStenCompute2DXYnp(&xyDirCompute, HOST);
    checkError("compute 2");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");

    FILE* f = std::fopen(argv[6], "wb");
    if (!f) {
        std::perror("fopen");
        return 2;
    }
    // the second apply wrote into the original input buffer
    std::fwrite(dataInput, sizeof(double), n, f);
    writeHandle(f, xyDirCompute, dataOutput);
    std::fclose(f);

    if (periodic) cuStenDestroy2DXYp(&xyDirCompute);
    else cuStenDestroy2DXYnp(&xyDirCompute);
    checkError("destroy");
    (void)hipFree(dataInput);
    (void)hipFree(dataOutput);
    (void)hipFree(weights);
    std::printf("weights_caller: %dx%d numTiles %d periodic %d done\n", nx, ny, numTiles,
                periodic);
    return 0;
}
