/*
 * custen_pent.h -- batched pentadiagonal systems on the GPU.
 *
 * Replaces launches of the reference's batch kernels
 *   pentFactorBatch<<<grid, block>>>(ds, dl, d, du, dw, m, batchCount)
 *   pentSolveBatch<<<grid, block>>>(ds, dl, d, du, dw, b, m, batchCount)
 * declared in cuPentCahnADI/src/cuPentBatch.h:28-52 and defined in
 * cuPentCahnADI/src/cuPentBatch.cu:35-198, with host entry points of the same
 * argument order plus a stream.
 *
 * Layout (as the reference): batchCount systems of m unknowns, interleaved --
 * element i of system k at index i * batchCount + k of each array.  Each
 * system has its own diagonals: ds (2 below the main diagonal; first two
 * entries unused), dl (1 below; first unused), d (main), du (1 above; last
 * unused), dw (2 above; last two unused).  All arrays are device memory of the
 * current device.  m >= 4.
 *
 * Both calls are asynchronous on `stream` (0: the null stream) and return
 * CUSTEN_SUCCESS or an error code (also recorded for custenGetLastError).
 */
#ifndef CUSTEN_PENT_H
#define CUSTEN_PENT_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-place LU factorisation without pivoting (pentFactorBatch): dl, d, du, dw
 * are overwritten with the factors; ds is read only. */
int custenPentFactorBatch(double* ds, double* dl, double* d, double* du, double* dw, int m,
                          int batchCount, hipStream_t stream);

/* Forward and back substitution with factored diagonals (pentSolveBatch):
 * b holds the right-hand sides on entry and the solutions on return. */
int custenPentSolveBatch(const double* ds, const double* dl, const double* d, const double* du,
                         const double* dw, double* b, int m, int batchCount, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUSTEN_PENT_H */
