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
/* Tile configuration of the fused steps (results do not depend on it):
 * band = 0: strip kernel, rowsPerWave 4 or 8 and chunks >= 1 row chunks per
 * workgroup; band = 1: band walker for the interior 128-column strips of
 * every grid (chunks per band), strip kernel for the rest.  Replaces the
 * autotuner's choice; graphs captured before are discarded.  Waits for the
 * batch's earlier work on its own stream. */
int custenBatchSetTileConfig(custenBatch* batch, int rowsPerWave, int chunks, int band);
/* The configuration in use: *band is 1 for the band walker.  All zero before
 * the first fused run when neither the autotuner nor the caller chose one. */
int custenBatchGetTileConfig(const custenBatch* batch, int* rowsPerWave, int* chunks, int* band);
void custenBatchDestroy(custenBatch* batch);

#ifdef __cplusplus
}
#endif

#endif /* CUSTEN_BATCH_H */
