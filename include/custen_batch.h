/*
 * custen_batch.h -- batched uneven grids (extension of the cuSten C ABI; no
 * counterpart in the reference, whose handles cover one grid each).
 *
 * Many independent periodic xy grids share one stencil (3x3 or 5x5 weights)
 * and are advanced together: each step is one launch over every grid, and a
 * run of steps is captured once into a HIP graph.  Per grid the results are
 * bit-identical to cuStenCompute2DXYp followed by cuStenSwap2DXYp.
 */
#ifndef CUSTEN_BATCH_H
#define CUSTEN_BATCH_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct custenBatch custenBatch;
/* dtype: 0 = FP64 storage, 1 = FP32 storage (then in/out hold float*). */
custenBatch* custenBatchCreate(int deviceNum, int numGrids, const int* nx, const int* ny,
                               void* const* dataOutput, void* const* dataInput,
                               const double* weights, int numStenHoriz, int numStenLeft,
                               int numStenVert, int numStenTop, int dtype);
/* Apply `steps` times, swapping input and output after every apply (as
 * cuStenSwap does).  mode 0 = one launch per step for all grids (captured in
 * a HIP graph on first use), mode 1 = plain per-grid launch loop. */
int custenBatchRun(custenBatch* batch, int steps, int mode, hipStream_t stream);
/* Where the current field of every grid is: 0 = dataInput[g] (before the
 * first step and after an even number of steps), 1 = dataOutput[g]. */
int custenBatchParity(const custenBatch* batch);
void custenBatchDestroy(custenBatch* batch);

#ifdef __cplusplus
}
#endif

#endif /* CUSTEN_BATCH_H */
