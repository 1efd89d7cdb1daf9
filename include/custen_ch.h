/*
 * custen_ch.h -- Cahn-Hilliard ADI solver and request server (extension of
 * the cuSten C ABI; the reference ships this as an application,
 * cuPentCahnADI/src/cuPentCahnADI.cu, not as a library API).
 *
 * Scheme per step (cuPentCahnADI.cu:528-596; serialCahnADI.c:1010-1047):
 *   cBar = 2c - cOld;  X = L5(cBar) + ((-2/3)(c - cOld) + L3(c^3 - c));
 *   cyclic pentadiagonal solves of X along x, then along y;
 *   cOld <- c;  c <- cBar + X.
 */
#ifndef CUSTEN_CH_H
#define CUSTEN_CH_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct custenCH custenCH;

typedef struct {
    int n;              /* grid n x n, periodic, n >= 8                      */
    double lx;          /* domain length (reference default 2 pi)            */
    double D;           /* mobility (1.0)                                    */
    double gamma;       /* interface parameter (0.01)                        */
    double dtOverDx;    /* dt = dtOverDx * dx (0.1)                          */
} custenCHParams;

/* Single-simulation solver with resident fields.  c0: n*n initial condition
 * (host or device); cOld starts equal to c0 as in the reference. */
custenCH* custenCHCreate(int deviceNum, const custenCHParams* params, const double* c0);
/* Advance `steps` steps.  mode 0: the fused step kernels are queued back to
 * back (four per step: two column solves, the transpose, and the final phase
 * fused with the next step's right-hand side); mode 1: the reference's kernel sequence with a stream
 * synchronisation after each launch (the reference's launch-and-synchronise loop, cuPentCahnADI.cu:528-596),
 * used as the plain per-grid baseline. */
int custenCHRun(custenCH* ch, int steps, int mode, hipStream_t stream);
/* Copy the current field (n*n doubles) to host memory. */
int custenCHGetField(custenCH* ch, double* c, double* cOld);
void custenCHDestroy(custenCH* ch);

/* Steps taken so far (including those recorded in a restored checkpoint). */
int custenCHStepsDone(custenCH* ch);

/* Snapshot output, the counterpart of Print_Out (cuPentCahnADI.cu:103-140,
 * which writes c to an HDF5 file per output time): the current field c as a
 * NumPy .npy file (float64, shape (n, n), C order).  Written to a temporary
 * name and renamed, so a reader never sees a partial file. */
int custenCHSnapshot(custenCH* ch, const char* path);

/* Checkpoint / restart (no counterpart in the reference): the parameters, the
 * step count and both time levels c and cOld, so that a restored solver
 * continues bit for bit as the original would have.  Format: "CUSTENCH",
 * u32 version (1), u32 n, f64 lx, D, gamma, dtOverDx, i64 steps done, then c
 * and cOld as n*n little-endian doubles each. */
int custenCHCheckpoint(custenCH* ch, const char* path);
custenCH* custenCHRestore(int deviceNum, const char* path);

/* Request server with continuous batching.  Requests are admitted (FIFO)
 * while their resident footprint fits in `hbmBudgetBytes` and a slot is free.
 * Requests of one grid size share a lane (one stream) whose batched launches
 * advance all of them together, up to 8 steps per batch and two batches in
 * flight; lanes of different sizes run concurrently.  Finished requests
 * retire and free their slot; their results are copied to host memory
 * asynchronously. */
typedef struct custenCHServer custenCHServer;
custenCHServer* custenCHServerCreate(int deviceNum, size_t hbmBudgetBytes, int maxSlots);
/* Queue a request.  Returns its id (>= 0) or -1.  c0 is uploaded to the
 * request's device buffers at submit when the budget has room (otherwise
 * copied to host memory until admission); the caller may reuse c0 as soon as
 * Submit returns. */
int custenCHServerSubmit(custenCHServer* s, const custenCHParams* params, int steps,
                         const double* c0);
/* The same for a simulation resumed from a checkpoint: both time levels. */
int custenCHServerSubmitResume(custenCHServer* s, const custenCHParams* params, int steps,
                               const double* c, const double* cOld);
/* Run engine iterations until the queue is empty and all requests finished
 * (maxIterations <= 0: no limit).  Returns the number of iterations. */
int custenCHServerRunUntilIdle(custenCHServer* s, int maxIterations);
/* One engine iteration (admit, step all, retire).  Returns the number of
 * requests still queued or resident. */
int custenCHServerStep(custenCHServer* s);
/* Request state: 0 queued, 1 resident, 2 finished.  On finish, latency in
 * seconds from submit and from admission are available. */
int custenCHServerQuery(custenCHServer* s, int id, int* state, double* latencySubmit,
                        double* latencyAdmit);
/* Copy a finished request's final field to host memory (n*n doubles). */
int custenCHServerResult(custenCHServer* s, int id, double* c);
/* Allocation-free advance.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Advances the simulation held in c (current field) and cOld
 * (previous field), both n*n doubles in device memory of deviceNum, by
 * `steps` fused steps on `stream`, with the same results as custenCHRun mode 0
 * (cOld = c at the start of a simulation, as in the reference).  On return
 * (asynchronously, in stream order) c holds the new current field and cOld the
 * previous one.  The work buffers (two n*n fields, the factored diagonals and
 * the descriptor, 16.1 MiB) are a static device array of the library, one per
 * GPU, held until the process ends; n is limited to [8, 1024] by its size and
 * larger n is refused with CUSTEN_ERROR_INVALID_VALUE.  The call allocates
 * nothing and copies nothing from the host.  Every call on one GPU shares that
 * array: calls that overlap on one GPU race and must be serialised by the
 * caller.  Returns CUSTEN_SUCCESS or an error code (also recorded for
 * custenGetLastError). */
int custenCHAdvanceNoAlloc(int deviceNum, const custenCHParams* params, double* c, double* cOld,
                           int steps, hipStream_t stream);

/* Bytes a request of grid size n keeps resident: five n*n fields (three time
 * levels used in turn and two work fields) and the factored matrix. */
size_t custenCHFootprint(int n);
void custenCHServerDestroy(custenCHServer* s);

#ifdef __cplusplus
}
#endif

#endif /* CUSTEN_CH_H */
