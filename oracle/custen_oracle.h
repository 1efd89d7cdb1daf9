/*
 * custen_oracle.h -- CPU oracle for the cuSten stencil path and the
 * Cahn-Hilliard ADI step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by the tests (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, explicit fma()).
 *
 * Every function states the reference file:line whose behaviour it restates.
 */
#ifndef CUSTEN_ORACLE_H
#define CUSTEN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stencils with weights ------------------------------------------------ */

/* Generic 2D apply covering all six weight variants.
 *   x  variants: V = 1, T = 0;   y variants: H = 1, L = 0.
 *   periodic != 0 : every point written, indices wrap (mod nx / mod ny).
 *   periodic == 0 : only points whose whole window is in the domain are
 *                   written; if zero_right_edge, columns [nx-R, nx) are set to
 *                   0.0 (the x-np right-block behaviour).
 * nthreads > 1 parallelises over rows with OpenMP (results are identical). */
void cso_apply_f64(const double* in, double* out, long nx, long ny,
                   const double* w, int H, int L, int V, int T,
                   int periodic, int zero_right_edge, int nthreads);
void cso_apply_f32(const float* in, float* out, long nx, long ny,
                   const float* w, int H, int L, int V, int T,
                   int periodic, int zero_right_edge, int nthreads);

/* ---- stencils with a point function ("Fun" variants) ----------------------- */

enum {
    CSO_FUN_X_CENTRAL_DIFF = 1,  /* (d[l-1] - 2 d[l] + d[l+1]) * coe[0]          */
    CSO_FUN_Y_WSUM         = 2,  /* sum_k coe[k] d[l + (k-T) jump], k < numCoe    */
    CSO_FUN_XY_WSUM        = 3,  /* sum_{j,i} coe[jH+i] d[l + j jump + i]         */
    CSO_FUN_XY_CUBE_MINUS  = 4,  /* sum_{j,i} coe[jH+i] (u^3 - u)                 */
    CSO_FUN_X_WSUM         = 5   /* sum_k coe[k] d[l - L + k], k < numCoe         */
};

/* dir: 0 = x, 1 = y, 2 = xy.  Write masks as for the weight variants, except
 * that no Fun variant zeroes the right edge. */
void cso_apply_fun_f64(int fun_id, int dir, const double* in, double* out,
                       long nx, long ny, const double* coe, int numCoe,
                       int H, int L, int V, int T, int periodic, int nthreads);

/* ---- WENO5 upwind advection (xy periodic) ---------------------------------- */

/* out = u * Fx + v * Fy with Fx, Fy the fifth-order WENO approximations of
 * d(in)/dx, d(in)/dy, upwinded on the sign of u (resp. v), periodic in x and
 * y.  coeDx = 1/dx, coeDy = 1/dy.  single_squares != 0 squares the smoothness
 * terms in single precision (the reference evaluates them with powf on
 * doubles); 0 squares them in double. */
void cso_weno_f64(const double* in, const double* u, const double* v, double* out,
                  long nx, long ny, double coeDx, double coeDy, int single_squares,
                  int nthreads);

/* ---- Cahn-Hilliard ADI ----------------------------------------------------- */

typedef struct {
    int n;              /* grid is n x n, periodic                               */
    int m;              /* n - 2, size of the reduced (non-cyclic) system        */
    double dx, dt, sigma_lin, sigma_nl;
    double a, b, c, d, e;
    double omega[4];
    double wlin[25];    /* 5x5 biharmonic RHS weights, row-major                 */
    double wnl[9];      /* 3x3 Laplacian weights of the nonlinear term           */
    double *ds, *dl, *diag, *du, *dw;   /* factored diagonals, length m          */
    double *inv1, *inv2;                /* E^-1 f1, E^-1 f2, length m            */
} cso_ch_t;

int  cso_ch_init(cso_ch_t* ch, int n, double lx, double D, double gamma,
                 double dt_over_dx);
void cso_ch_free(cso_ch_t* ch);
/* One ADI step.  c, cOld: n*n fields (updated in place).  work: 3*n*n doubles. */
void cso_ch_step(const cso_ch_t* ch, double* c, double* cOld, double* work);
/* Solve the cyclic pentadiagonal system in place for one line of length n with
 * element stride `stride`. */
void cso_ch_cyclic_solve(const cso_ch_t* ch, double* x, long stride);

/* pentFactor / pentSolve on explicit diagonals (length m, in place). */
void cso_pent_factor(double* ds, double* dl, double* d, double* du, double* dw, int m);
void cso_pent_solve(const double* ds, const double* dl, const double* d,
                    const double* du, const double* dw, double* b, int m);
/* The same on `batch` interleaved systems (element i of system k at
 * i * batch + k), as cuPentBatch's pentFactorBatch / pentSolveBatch. */
void cso_pent_factor_batch(double* ds, double* dl, double* d, double* du, double* dw, int m,
                           long batch, int nthreads);
void cso_pent_solve_batch(const double* ds, const double* dl, const double* d,
                          const double* du, const double* dw, double* b, int m, long batch,
                          int nthreads);

/* ---- synthetic inputs ------------------------------------------------------- */

/* SplitMix64 stream: element i uses state seed + (i+1)*0x9E3779B97F4A7C15.
 * u = (z >> 11) * 2^-53, value = a + (b - a) * u. */
void cso_fill_uniform(double* out, size_t n, uint64_t seed, double a, double b);

#ifdef __cplusplus
}
#endif
#endif
