// pent_batch.hip -- batched pentadiagonal factorisation and solve on
// interleaved systems (cuPentCahnADI/src/cuPentBatch.cu:35-198): element i of
// system k lives at [i * batch + k], every system has its own diagonals.
//
// One lane per system, so each row access of a wave is one contiguous
// 512-byte segment per array.  The recurrences are sequential in the row, so
// a lane keeps the previous rows' results in registers and prefetches the
// next kPF rows of every array it reads while it computes the current kPF
// (double-buffered), which keeps enough bytes in flight even when a batch
// fills only one wave per SIMD.  Arithmetic is the reference's, one IEEE
// operation at a time (contraction off), identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/custen_pent.h"
#include "plan_internal.h"

#pragma clang fp contract(off)

namespace custen {
namespace pentdev {

constexpr int kPF = 8;
constexpr int kBlock = 64;

// pentFactorBatch (cuPentBatch.cu:35-113)
__global__ __launch_bounds__(kBlock) void factorKernel(double* __restrict__ ds,
                                                       double* __restrict__ dl,
                                                       double* __restrict__ d,
                                                       double* __restrict__ du,
                                                       double* __restrict__ dw, int m,
                                                       int64_t batch)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= batch) return;
    ds += k;
    dl += k;
    d += k;
    du += k;
    dw += k;
    const int64_t s = batch;
    // row 0
    const double d0 = d[0];
    const double du0 = du[0] / d0;
    const double dw0 = dw[0] / d0;
    du[0] = du0;
    dw[0] = dw0;
    // row 1
    const double dl1 = dl[s];
    const double d1 = d[s] - dl1 * du0;
    const double du1 = (du[s] - dl1 * dw0) / d1;
    const double dw1 = dw[s] / d1;
    d[s] = d1;
    du[s] = du1;
    dw[s] = dw1;
    double duPP = du0, dwPP = dw0, duP = du1, dwP = dw1;

    double cs[kPF], cl[kPF], cd[kPF], cu[kPF], cw[kPF];
    double ns[kPF], nl[kPF], nd[kPF], nu[kPF], nw[kPF];
#pragma unroll
    for (int j = 0; j < kPF; j++) {
        const int i = 2 + j;
        if (i < m) {
            const int64_t o = (int64_t)i * s;
            cs[j] = ds[o];
            cl[j] = dl[o];
            cd[j] = d[o];
            cu[j] = du[o];
            cw[j] = dw[o];
        }
    }
    for (int i0 = 2; i0 < m; i0 += kPF) {
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 + kPF + j;
            if (i < m) {
                const int64_t o = (int64_t)i * s;
                ns[j] = ds[o];
                nl[j] = dl[o];
                nd[j] = d[o];
                nu[j] = du[o];
                nw[j] = dw[o];
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 + j;
            if (i < m) {
                const int64_t o = (int64_t)i * s;
                const double lv = cl[j] - cs[j] * duPP;
                const double dv = (cd[j] - cs[j] * dwPP) - lv * duP;
                dl[o] = lv;
                d[o] = dv;
                double uv = 0.0, wv = cw[j];
                if (i < m - 1) {  // rows up to m-2 update du
                    uv = (cu[j] - lv * dwP) / dv;
                    du[o] = uv;
                }
                if (i < m - 2) {  // interior rows also update dw
                    wv = cw[j] / dv;
                    dw[o] = wv;
                }
                duPP = duP;
                dwPP = dwP;
                duP = uv;
                dwP = wv;
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            cs[j] = ns[j];
            cl[j] = nl[j];
            cd[j] = nd[j];
            cu[j] = nu[j];
            cw[j] = nw[j];
        }
    }
}

// pentSolveBatch (cuPentBatch.cu:119-198)
__global__ __launch_bounds__(kBlock) void solveKernel(const double* __restrict__ ds,
                                                      const double* __restrict__ dl,
                                                      const double* __restrict__ d,
                                                      const double* __restrict__ du,
                                                      const double* __restrict__ dw,
                                                      double* __restrict__ b, int m, int64_t batch)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= batch) return;
    ds += k;
    dl += k;
    d += k;
    du += k;
    dw += k;
    b += k;
    const int64_t s = batch;

    // forward substitution
    double b2 = b[0] / d[0];
    b[0] = b2;
    double b1 = (b[s] - dl[s] * b2) / d[s];
    b[s] = b1;
    double cs[kPF], cl[kPF], cd[kPF], cb[kPF];
    double ns[kPF], nl[kPF], nd[kPF], nb[kPF];
#pragma unroll
    for (int j = 0; j < kPF; j++) {
        const int i = 2 + j;
        if (i < m) {
            const int64_t o = (int64_t)i * s;
            cs[j] = ds[o];
            cl[j] = dl[o];
            cd[j] = d[o];
            cb[j] = b[o];
        }
    }
    for (int i0 = 2; i0 < m; i0 += kPF) {
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 + kPF + j;
            if (i < m) {
                const int64_t o = (int64_t)i * s;
                ns[j] = ds[o];
                nl[j] = dl[o];
                nd[j] = d[o];
                nb[j] = b[o];
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 + j;
            if (i < m) {
                const double v = ((cb[j] - cs[j] * b2) - cl[j] * b1) / cd[j];
                b[(int64_t)i * s] = v;
                b2 = b1;
                b1 = v;
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            cs[j] = ns[j];
            cl[j] = nl[j];
            cd[j] = nd[j];
            cb[j] = nb[j];
        }
    }

    // backward substitution: x[m-1] = b1, x[m-2] = b2 - du[m-2] x[m-1]
    double a1 = b1;
    double a0 = b2 - du[(int64_t)(m - 2) * s] * a1;
    b[(int64_t)(m - 2) * s] = a0;
    double cuu[kPF], cww[kPF], nuu[kPF], nww[kPF];
#pragma unroll
    for (int j = 0; j < kPF; j++) {
        const int i = m - 3 - j;
        if (i >= 0) {
            const int64_t o = (int64_t)i * s;
            cuu[j] = du[o];
            cww[j] = dw[o];
            cb[j] = b[o];
        }
    }
    for (int i0 = m - 3; i0 >= 0; i0 -= kPF) {
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 - kPF - j;
            if (i >= 0) {
                const int64_t o = (int64_t)i * s;
                nuu[j] = du[o];
                nww[j] = dw[o];
                nb[j] = b[o];
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            const int i = i0 - j;
            if (i >= 0) {
                const double v = (cb[j] - cuu[j] * a0) - cww[j] * a1;
                b[(int64_t)i * s] = v;
                a1 = a0;
                a0 = v;
            }
        }
#pragma unroll
        for (int j = 0; j < kPF; j++) {
            cuu[j] = nuu[j];
            cww[j] = nww[j];
            cb[j] = nb[j];
        }
    }
}

}  // namespace pentdev

namespace {

int check(const char* fn, const void* const* ptrs, int nptr, int m, int batchCount)
{
    if (m < 4) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": m must be >= 4");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    if (batchCount < 1) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": batchCount must be >= 1");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return lastError();
    }
    for (int i = 0; i < nptr; i++) {
        if (ptrs[i] == nullptr) {
            setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": null array");
            return CUSTEN_ERROR_INVALID_VALUE;
        }
        if (classify(ptrs[i], dev) != MemKind::Device) {
            setError(CUSTEN_ERROR_INVALID_VALUE,
                     std::string(fn) + ": arrays must be device memory of the current device");
            return CUSTEN_ERROR_INVALID_VALUE;
        }
    }
    return CUSTEN_SUCCESS;
}

}  // namespace
}  // namespace custen

using namespace custen;

extern "C" int custenPentFactorBatch(double* ds, double* dl, double* d, double* du, double* dw,
                                     int m, int batchCount, hipStream_t stream)
{
    const void* ptrs[] = {ds, dl, d, du, dw};
    int rc = check("custenPentFactorBatch", ptrs, 5, m, batchCount);
    if (rc != CUSTEN_SUCCESS) return rc;
    const unsigned blocks = (unsigned)((batchCount + pentdev::kBlock - 1) / pentdev::kBlock);
    hipLaunchKernelGGL(pentdev::factorKernel, dim3(blocks), dim3(pentdev::kBlock), 0, stream, ds,
                       dl, d, du, dw, m, (int64_t)batchCount);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        setHipError(e, "custenPentFactorBatch");
        return lastError();
    }
    return CUSTEN_SUCCESS;
}

extern "C" int custenPentSolveBatch(const double* ds, const double* dl, const double* d,
                                    const double* du, const double* dw, double* b, int m,
                                    int batchCount, hipStream_t stream)
{
    const void* ptrs[] = {ds, dl, d, du, dw, b};
    int rc = check("custenPentSolveBatch", ptrs, 6, m, batchCount);
    if (rc != CUSTEN_SUCCESS) return rc;
    const unsigned blocks = (unsigned)((batchCount + pentdev::kBlock - 1) / pentdev::kBlock);
    hipLaunchKernelGGL(pentdev::solveKernel, dim3(blocks), dim3(pentdev::kBlock), 0, stream, ds,
                       dl, d, du, dw, b, m, (int64_t)batchCount);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        setHipError(e, "custenPentSolveBatch");
        return lastError();
    }
    return CUSTEN_SUCCESS;
}
