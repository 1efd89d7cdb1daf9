/*
 * custen_oracle.c -- CPU oracle for the cuSten stencil path and the
 * Cahn-Hilliard ADI step.
 *
 * TEST INFRASTRUCTURE ONLY: the tests, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this as the checker.  The product library never links
 * or calls it.
 *
 * Arithmetic contract (shared with the HIP kernels, which are compiled with
 * -ffp-contract=off and spell every fused multiply-add explicitly):
 *   - stencil sums are an fma chain starting from 0.0, j (rows) outer,
 *     i (columns) inner, weights in row-major order;
 *   - everything else is evaluated operation by operation without contraction,
 *     in the order the reference's serial C code writes it.
 * Built with -ffp-contract=off so the compiler adds no contraction of its own.
 */
#include "custen_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline long wrap_index(long i, long n)
{
    long r = i % n;
    return r < 0 ? r + n : r;
}

/* --------------------------------------------------------------------------
 * Weight stencils.
 *
 * Summation: cuSten/src/kernels/2d_xy_p_kernel.cu:505-520 (j outer, i inner,
 * sum starts at 0.0, `sum += w * u` contracted to an fma by nvcc);
 * 2d_x_p_kernel.cu:166-172 (x: k ascending from x-L); 2d_y_p_kernel.cu:121-127
 * (y: k ascending from y-T).  The x and y variants are the xy sum with V = 1 or
 * H = 1, which gives the same term order.
 *
 * Write masks:
 *   periodic: every point (2d_xy_p_kernel.cu:527-529).
 *   np: x in [L, nx-R), y in [T, ny-B) (2d_xy_np_kernel.cu:208-221, 784-787;
 *       the right-edge test at :862 uses numStenLeft, which equals numStenRight
 *       for symmetric stencils; this restatement uses numStenRight).
 *   x-np additionally writes 0.0 at x in [nx-R, nx)
 *       (2d_x_np_kernel.cu:164-176: `sum` stays 0.0 for lanes >= BLOCK_X - R).
 * -------------------------------------------------------------------------- */

#define DEFINE_APPLY(NAME, T_, FMA)                                                      \
void NAME(const T_* in, T_* out, long nx, long ny, const T_* w, int H, int L, int V,     \
          int T, int periodic, int zero_right_edge, int nthreads)                       \
{                                                                                       \
    const int R = H - 1 - L;                                                            \
    const int B = V - 1 - T;                                                            \
    long ylo = 0, yhi = ny, xlo = 0, xhi = nx;                                          \
    if (!periodic) { ylo = T; yhi = ny - B; xlo = L; xhi = nx - R; }                    \
    if (nthreads < 1) nthreads = 1;                                                     \
    _Pragma("omp parallel num_threads(nthreads)")                                       \
    {                                                                                   \
        T_* ext = (T_*)malloc(sizeof(T_) * (size_t)V * (size_t)(nx + H));               \
        T_* acc = (T_*)malloc(sizeof(T_) * (size_t)nx);                                 \
        _Pragma("omp for schedule(static)")                                             \
        for (long y = 0; y < ny; y++) {                                                 \
            if (!periodic && zero_right_edge && V == 1) {                               \
                for (long x = (nx - R > 0 ? nx - R : 0); x < nx; x++)                   \
                    out[y * nx + x] = (T_)0.0;                                          \
            }                                                                           \
            if (y < ylo || y >= yhi || xlo >= xhi) continue;                            \
            /* halo-extended copies of the V input rows: ext[j][k] = u(y-T+j, k-L) */   \
            for (int j = 0; j < V; j++) {                                               \
                long yy = y - T + j;                                                    \
                if (periodic) yy = wrap_index(yy, ny);                                  \
                const T_* row = in + yy * nx;                                           \
                T_* e = ext + (size_t)j * (size_t)(nx + H);                             \
                for (long k = 0; k < nx + H - 1; k++) {                                 \
                    long xx = k - L;                                                    \
                    if (periodic) xx = wrap_index(xx, nx);                              \
                    else if (xx < 0 || xx >= nx) { e[k] = (T_)0.0; continue; }          \
                    e[k] = row[xx];                                                     \
                }                                                                       \
            }                                                                           \
            for (long x = xlo; x < xhi; x++) acc[x] = (T_)0.0;                          \
            for (int j = 0; j < V; j++) {                                               \
                const T_* e = ext + (size_t)j * (size_t)(nx + H);                       \
                for (int i = 0; i < H; i++) {                                           \
                    const T_ wk = w[j * H + i];                                         \
                    for (long x = xlo; x < xhi; x++)                                    \
                        acc[x] = FMA(wk, e[x + i], acc[x]);                             \
                }                                                                       \
            }                                                                           \
            for (long x = xlo; x < xhi; x++) out[y * nx + x] = acc[x];                  \
        }                                                                               \
        free(ext);                                                                      \
        free(acc);                                                                      \
    }                                                                                   \
}

DEFINE_APPLY(cso_apply_f64, double, fma)
DEFINE_APPLY(cso_apply_f32, float, fmaf)

/* --------------------------------------------------------------------------
 * Point-function ("Fun") stencils.
 *
 * The user function sees a local copy of the window:
 *   x : f(data, coe, loc)            loc = window centre
 *       (2d_x_np_fun_kernel.cu:119-121, 137-140, 162-165)
 *   y : f(data, coe, loc, jump)      loc = window centre
 *       (2d_y_p_fun_kernel.cu:133-138)
 *   xy: f(data, coe, loc, jump, H, V) loc = window top-left
 *       (2d_xy_p_fun_kernel.cu:516-526)
 * Only the values at the offsets the function reads matter, so the oracle
 * passes a compact window with jump = H.
 *
 * The functions below restate the reference's example point functions:
 *   X_CENTRAL_DIFF : examples/src/2d_x_np_fun.cu:49-54
 *   Y_WSUM         : examples/src/2d_y_p_fun.cu:49-62 (9 coefficients, loc-4*jump)
 *                    generalised to numCoe = V coefficients starting at loc-T*jump
 *   XY_WSUM        : examples/src/2d_xy_p_fun.cu:49-68
 *   XY_CUBE_MINUS  : cuPentCahnADI/src/cuPentCahnADI.cu:164-188 (nonLinRHS)
 *                    with u^3-u evaluated as (u*u)*u - u, unfused, as
 *                    serialCahnADI.c:617 does, and the sum as an fma chain.
 *   X_WSUM         : numCoe = H coefficients, sum_k coe[k] d[loc-L+k]
 * -------------------------------------------------------------------------- */

static double fun_eval(int fun_id, const double* data, const double* coe, int loc,
                       int jump, int H, int V, int L, int T, int numCoe)
{
    double r = 0.0;
    switch (fun_id) {
    case CSO_FUN_X_CENTRAL_DIFF: {
        double t = data[loc - 1] - 2.0 * data[loc];
        t = t + data[loc + 1];
        return t * coe[0];
    }
    case CSO_FUN_X_WSUM:
        for (int k = 0; k < numCoe; k++) r = fma(coe[k], data[loc - L + k], r);
        return r;
    case CSO_FUN_Y_WSUM:
        for (int k = 0; k < numCoe; k++) r = fma(coe[k], data[loc + (k - T) * jump], r);
        return r;
    case CSO_FUN_XY_WSUM: {
        int cnt = 0;
        for (int j = 0; j < V; j++)
            for (int i = 0; i < H; i++) r = fma(coe[cnt++], data[loc + j * jump + i], r);
        return r;
    }
    case CSO_FUN_XY_CUBE_MINUS: {
        int cnt = 0;
        for (int j = 0; j < V; j++)
            for (int i = 0; i < H; i++) {
                double u = data[loc + j * jump + i];
                double u2 = u * u;
                double u3 = u2 * u;
                double q = u3 - u;
                r = fma(coe[cnt++], q, r);
            }
        return r;
    }
    default:
        return NAN;
    }
}

void cso_apply_fun_f64(int fun_id, int dir, const double* in, double* out,
                       long nx, long ny, const double* coe, int numCoe,
                       int H, int L, int V, int T, int periodic, int nthreads)
{
    const int R = H - 1 - L;
    const int B = V - 1 - T;
    long ylo = 0, yhi = ny, xlo = 0, xhi = nx;
    if (!periodic) { ylo = T; yhi = ny - B; xlo = L; xhi = nx - R; }
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long y = ylo; y < yhi; y++) {
        double win[64 * 64];
        for (long x = xlo; x < xhi; x++) {
            for (int j = 0; j < V; j++) {
                long yy = y - T + j;
                if (periodic) yy = wrap_index(yy, ny);
                for (int i = 0; i < H; i++) {
                    long xx = x - L + i;
                    if (periodic) xx = wrap_index(xx, nx);
                    win[j * H + i] = in[yy * nx + xx];
                }
            }
            int loc;
            if (dir == 0) loc = L;               /* centre of a 1 x H window */
            else if (dir == 1) loc = T * H;      /* centre of a V x 1 window */
            else loc = 0;                        /* top-left                  */
            out[y * nx + x] = fun_eval(fun_id, win, coe, loc, H, H, V, L, T, numCoe);
        }
    }
}

/* --------------------------------------------------------------------------
 * WENO5 upwind advection, xy periodic.
 *
 * Restates cuSten/src/kernels/2d_xyADVWENO_p_kernel.cu:
 *   wenoSten :51-86 (Fedkiw's notation: three candidate stencils phi1..3,
 *            smoothness s1..s3 with the squares taken by powf on doubles,
 *            weights 0.1 / 0.6 / 0.3 over (s + 1e-6)^2, normalised by a
 *            reciprocal), differences (q[k] - q[k-1]) * coeD, upwind choice on
 *            u > 0 (else the mirrored five differences) :284-346 for x and
 *            :348-397 for y, output u * Fx + v * Fy :401.
 * Every operation is a separate IEEE double operation in source order; the
 * single-precision squares convert to float, square in float and widen back.
 * -------------------------------------------------------------------------- */
static double weno_sq(double d, int single_squares)
{
    if (single_squares) {
        float f = (float)d;
        float f2 = f * f;
        return (double)f2;
    }
    return d * d;
}

static double weno_flux(double v1, double v2, double v3, double v4, double v5, int ss)
{
    const double phi1 = ((1.0 / 3.0) * v1 - (7.0 / 6.0) * v2) + (11.0 / 6.0) * v3;
    const double phi2 = (-(1.0 / 6.0) * v2 + (5.0 / 6.0) * v3) + (1.0 / 3.0) * v4;
    const double phi3 = ((1.0 / 3.0) * v3 + (5.0 / 6.0) * v4) - (1.0 / 6.0) * v5;
    const double s1 = (13.0 / 12.0) * weno_sq((v1 - 2.0 * v2) + v3, ss) +
                      0.25 * weno_sq((v1 - 4.0 * v2) + 3.0 * v3, ss);
    const double s2 = (13.0 / 12.0) * weno_sq((v2 - 2.0 * v3) + v4, ss) +
                      0.25 * weno_sq(v2 - v4, ss);
    const double s3 = (13.0 / 12.0) * weno_sq((v3 - 2.0 * v4) + v5, ss) +
                      0.25 * weno_sq((3.0 * v3 - 4.0 * v4) + v5, ss);
    const double a1 = 0.1 / weno_sq(s1 + 1e-06, ss);
    const double a2 = 0.6 / weno_sq(s2 + 1e-06, ss);
    const double a3 = 0.3 / weno_sq(s3 + 1e-06, ss);
    const double denom = 1.0 / ((a1 + a2) + a3);
    const double w1 = a1 * denom, w2 = a2 * denom, w3 = a3 * denom;
    return (phi1 * w1 + phi2 * w2) + phi3 * w3;
}

void cso_weno_f64(const double* in, const double* u, const double* v, double* out,
                  long nx, long ny, double coeDx, double coeDy, int single_squares,
                  int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long y = 0; y < ny; y++) {
        const double* row = in + y * nx;
        for (long x = 0; x < nx; x++) {
            double dx[7], dy[7];  /* d[k] = difference ending at offset k - 3, k = 1..6 */
            for (int k = 1; k < 7; k++) {
                const long xa = ((x + k - 3) % nx + nx) % nx, xb = ((x + k - 4) % nx + nx) % nx;
                const long ya = ((y + k - 3) % ny + ny) % ny, yb = ((y + k - 4) % ny + ny) % ny;
                dx[k] = (row[xa] - row[xb]) * coeDx;
                dy[k] = (in[ya * nx + x] - in[yb * nx + x]) * coeDy;
            }
            const long i = y * nx + x;
            double fx, fy;
            if (u[i] > 0.0) fx = weno_flux(dx[1], dx[2], dx[3], dx[4], dx[5], single_squares);
            else fx = weno_flux(dx[6], dx[5], dx[4], dx[3], dx[2], single_squares);
            if (v[i] > 0.0) fy = weno_flux(dy[1], dy[2], dy[3], dy[4], dy[5], single_squares);
            else fy = weno_flux(dy[6], dy[5], dy[4], dy[3], dy[2], single_squares);
            out[i] = u[i] * fx + v[i] * fy;
        }
    }
}

/* --------------------------------------------------------------------------
 * Cahn-Hilliard ADI.
 *
 * Restates cuPentSpeedUp/serialCahnADITiming/serialCahnADI.c:
 *   setLHS :40-84, pentFactor :91-155, pentSolve :161-241, setInv1/2 :247-319,
 *   findOmega :325-417, findCBar :423-437, findRHS :443-472,
 *   linearRHS :478-547, nonlinearRHS :553-622, solveEnd :628-655,
 *   solveFull :661-678, cyclicInv :684-712, transpose :718-727,
 *   findNew :733-747, parameters :776-945, time loop :1010-1047.
 * The two stencil sums are evaluated as the cuSten fma chain, because the GPU
 * program computes them with cuStenCompute2DXYp / cuStenCompute2DXYpFun
 * (cuPentCahnADI/src/cuPentCahnADI.cu:537-540).
 * -------------------------------------------------------------------------- */

static void pent_factor_strided(double* ds, double* dl, double* d, double* du, double* dw,
                                int m, long s)
{
    /* serialCahnADI.c:91-155; cuPentBatch.cu:35-113 is the same recurrence on
     * interleaved systems (element i of a system at i * s) */
    long r = 0;
    du[r] = du[r] / d[r];
    dw[r] = dw[r] / d[r];
    long rp = r;
    r += s;
    d[r] = d[r] - dl[r] * du[rp];
    du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    dw[r] = dw[r] / d[r];
    for (int i = 2; i < m - 2; i++) {
        long rpp = r - s;
        rp = r;
        r += s;
        dl[r] = dl[r] - ds[r] * du[rpp];
        d[r] = d[r] - ds[r] * dw[rpp] - dl[r] * du[rp];
        dw[r] = dw[r] / d[r];
        du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    }
    {
        long rpp = r - s;
        rp = r;
        r += s;
        dl[r] = dl[r] - ds[r] * du[rpp];
        d[r] = d[r] - ds[r] * dw[rpp] - dl[r] * du[rp];
        du[r] = (du[r] - dl[r] * dw[rp]) / d[r];
    }
    {
        long rpp = r - s;
        rp = r;
        r += s;
        dl[r] = dl[r] - ds[r] * du[rpp];
        d[r] = d[r] - ds[r] * dw[rpp] - dl[r] * du[rp];
    }
}

void cso_pent_factor(double* ds, double* dl, double* d, double* du, double* dw, int m)
{
    pent_factor_strided(ds, dl, d, du, dw, m, 1);
}

/* Diagonals with element stride ds_s (0 for diagonals shared by every line),
 * right-hand side with stride s. */
static void pent_solve_strided2(const double* ds, const double* dl, const double* d,
                                const double* du, const double* dw, long q, double* b, int m,
                                long s)
{
    /* serialCahnADI.c:161-241; cuPentBatch.cu:119-198 */
    b[0] = b[0] / d[0];
    b[s] = (b[s] - dl[q] * b[0]) / d[q];
    for (int i = 2; i < m; i++)
        b[i * s] = (b[i * s] - ds[i * q] * b[(i - 2) * s] - dl[i * q] * b[(i - 1) * s]) / d[i * q];
    b[(m - 2) * s] = b[(m - 2) * s] - du[(m - 2) * q] * b[(m - 1) * s];
    for (int i = m - 3; i >= 0; i--)
        b[i * s] = b[i * s] - du[i * q] * b[(i + 1) * s] - dw[i * q] * b[(i + 2) * s];
}

static void pent_solve_strided(const double* ds, const double* dl, const double* d,
                               const double* du, const double* dw, double* b, int m,
                               long s)
{
    pent_solve_strided2(ds, dl, d, du, dw, 1, b, m, s);
}

void cso_pent_factor_batch(double* ds, double* dl, double* d, double* du, double* dw, int m,
                           long batch, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long k = 0; k < batch; k++)
        pent_factor_strided(ds + k, dl + k, d + k, du + k, dw + k, m, batch);
}

void cso_pent_solve_batch(const double* ds, const double* dl, const double* d,
                          const double* du, const double* dw, double* b, int m, long batch,
                          int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long k = 0; k < batch; k++)
        pent_solve_strided2(ds + k, dl + k, d + k, du + k, dw + k, batch, b + k, m, batch);
}

void cso_pent_solve(const double* ds, const double* dl, const double* d,
                    const double* du, const double* dw, double* b, int m)
{
    pent_solve_strided(ds, dl, d, du, dw, b, m, 1);
}

void cso_ch_cyclic_solve(const cso_ch_t* ch, double* x, long s)
{
    const int n = ch->n, m = ch->m;
    pent_solve_strided(ch->ds, ch->dl, ch->diag, ch->du, ch->dw, x, m, s);
    /* solveEnd, serialCahnADI.c:628-655 */
    double nx2 = x[(n - 2) * s] - (ch->e * x[0] + ch->a * x[(n - 4) * s] + ch->b * x[(n - 3) * s]);
    double nx1 = x[(n - 1) * s] - (ch->d * x[0] + ch->e * x[s] + ch->a * x[(n - 3) * s]);
    x[(n - 2) * s] = ch->omega[0] * nx2 + ch->omega[1] * nx1;
    x[(n - 1) * s] = ch->omega[2] * nx2 + ch->omega[3] * nx1;
    /* solveFull, serialCahnADI.c:661-678 */
    double o2 = x[(n - 2) * s], o1 = x[(n - 1) * s];
    for (int i = 0; i < n - 2; i++)
        x[i * s] = x[i * s] - (ch->inv1[i] * o2 + ch->inv2[i] * o1);
}

int cso_ch_init(cso_ch_t* ch, int n, double lx, double D, double gamma, double dt_over_dx)
{
    memset(ch, 0, sizeof(*ch));
    if (n < 6) return -1;
    ch->n = n;
    ch->m = n - 2;
    const int m = ch->m;
    /* serialCahnADI.c:783-790, 913-921, 960 */
    ch->dx = lx / n;
    ch->dt = dt_over_dx * ch->dx;
    const double dx = ch->dx, dt = ch->dt;
    ch->sigma_lin = 2.0 * dt * D * gamma / (3.0 * (dx * dx * dx * dx));
    ch->sigma_nl = (dt / 3.0) * D * (2.0 / (dx * dx));
    const double s = ch->sigma_lin;
    ch->a = s;
    ch->b = -4 * s;
    ch->c = 1 + 6 * s;
    ch->d = -4 * s;
    ch->e = s;
    /* weights, serialCahnADI.c:947-951 and 968-970 */
    static const double lin_pat[25] = {0, 0, -1, 0, 0,   0, -2, 8, -2, 0,   -1, 8, -20, 8, -1,
                                       0, -2, 8, -2, 0,  0, 0, -1, 0, 0};
    for (int k = 0; k < 25; k++) ch->wlin[k] = lin_pat[k] * s;
    static const double nl_pat[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
    for (int k = 0; k < 9; k++) ch->wnl[k] = nl_pat[k] * ch->sigma_nl;

    ch->ds = calloc((size_t)m, sizeof(double));
    ch->dl = calloc((size_t)m, sizeof(double));
    ch->diag = calloc((size_t)m, sizeof(double));
    ch->du = calloc((size_t)m, sizeof(double));
    ch->dw = calloc((size_t)m, sizeof(double));
    ch->inv1 = calloc((size_t)m, sizeof(double));
    ch->inv2 = calloc((size_t)m, sizeof(double));
    if (!ch->ds || !ch->dl || !ch->diag || !ch->du || !ch->dw || !ch->inv1 || !ch->inv2) {
        cso_ch_free(ch);
        return -2;
    }
    /* setLHS, serialCahnADI.c:40-84 */
    for (int i = 0; i < m; i++) {
        if (i > 1) ch->ds[i] = ch->a;
        if (i > 0) ch->dl[i] = ch->b;
        ch->diag[i] = ch->c;
        if (i < m - 1) ch->du[i] = ch->d;
        if (i < m - 2) ch->dw[i] = ch->e;
    }
    cso_pent_factor(ch->ds, ch->dl, ch->diag, ch->du, ch->dw, m);

    /* findOmega, serialCahnADI.c:325-417: factor a second copy, solve E x = f1, f2 */
    double *fds = malloc(sizeof(double) * m), *fdl = malloc(sizeof(double) * m),
           *fd = malloc(sizeof(double) * m), *fdu = malloc(sizeof(double) * m),
           *fdw = malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) {
        fds[i] = ch->ds[i]; fdl[i] = ch->dl[i]; fd[i] = ch->diag[i];
        fdu[i] = ch->du[i]; fdw[i] = ch->dw[i];
    }
    /* setInv1 / setInv2, serialCahnADI.c:247-319 */
    for (int i = 0; i < m; i++) {
        ch->inv1[i] = 0.0;
        ch->inv2[i] = 0.0;
    }
    ch->inv1[0] = ch->a;
    ch->inv1[m - 2] = ch->e;
    ch->inv1[m - 1] = ch->d;
    ch->inv2[0] = ch->b;
    ch->inv2[1] = ch->a;
    ch->inv2[m - 1] = ch->e;
    cso_pent_solve(fds, fdl, fd, fdu, fdw, ch->inv1, m);
    cso_pent_solve(fds, fdl, fd, fdu, fdw, ch->inv2, m);
    free(fds); free(fdl); free(fd); free(fdu); free(fdw);
    const double a = ch->a, b = ch->b, c = ch->c, d = ch->d, e = ch->e;
    const double *i1 = ch->inv1, *i2 = ch->inv2;
    double z11 = e * i1[0] + a * i1[m - 2] + b * i1[m - 1];
    double z12 = e * i2[0] + a * i2[m - 2] + b * i2[m - 1];
    double z21 = d * i1[0] + e * i1[1] + a * i1[m - 1];
    double z22 = d * i2[0] + e * i2[1] + a * i2[m - 1];
    double y11 = c - z11, y12 = d - z12, y21 = b - z21, y22 = c - z22;
    double det = 1.0 / (y11 * y22 - y21 * y12);
    ch->omega[0] = y22 * det;
    ch->omega[1] = -y12 * det;
    ch->omega[2] = -y21 * det;
    ch->omega[3] = y11 * det;
    return 0;
}

void cso_ch_free(cso_ch_t* ch)
{
    free(ch->ds); free(ch->dl); free(ch->diag); free(ch->du); free(ch->dw);
    free(ch->inv1); free(ch->inv2);
    ch->ds = ch->dl = ch->diag = ch->du = ch->dw = ch->inv1 = ch->inv2 = NULL;
}

void cso_ch_step(const cso_ch_t* ch, double* c, double* cOld, double* work)
{
    const long n = ch->n, nn = n * n;
    double* cBar = work;
    double* cHalf = work + nn;
    double* nl = work + 2 * nn;
    /* findCBar, serialCahnADI.c:423-437 */
    for (long k = 0; k < nn; k++) cBar[k] = 2.0 * c[k] - cOld[k];
    /* linearRHS (5x5 on cBar) and nonlinearRHS (3x3 on c^3-c) */
    cso_apply_f64(cBar, cHalf, n, n, ch->wlin, 5, 2, 5, 2, 1, 0, 1);
    cso_apply_fun_f64(CSO_FUN_XY_CUBE_MINUS, 2, c, nl, n, n, ch->wnl, 9, 3, 1, 3, 1, 1, 1);
    /* findRHS, serialCahnADI.c:443-472 */
    const double k23 = -(2.0 / 3.0);
    for (long k = 0; k < nn; k++) {
        double t = k23 * (c[k] - cOld[k]) + nl[k];
        cHalf[k] = cHalf[k] + t;
        cOld[k] = c[k];
    }
    /* x-direction solves: each row is one system */
    for (long j = 0; j < n; j++) cso_ch_cyclic_solve(ch, cHalf + j * n, 1);
    /* y-direction solves: each column is one system (the serial code transposes,
       solves rows and transposes back; the arithmetic per system is identical) */
    for (long i = 0; i < n; i++) cso_ch_cyclic_solve(ch, cHalf + i, n);
    /* findNew, serialCahnADI.c:733-747 */
    for (long k = 0; k < nn; k++) c[k] = cBar[k] + cHalf[k];
}

/* -------------------------------------------------------------------------- */

void cso_fill_uniform(double* out, size_t n, uint64_t seed, double a, double b)
{
    const double span = b - a;
    for (size_t i = 0; i < n; i++) {
        uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        double u = (double)(z >> 11) * 0x1.0p-53;
        out[i] = a + span * u;
    }
}
