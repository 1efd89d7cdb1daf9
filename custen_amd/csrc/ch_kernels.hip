// ch_kernels.hip -- the Cahn-Hilliard ADI step on gfx950, batched over any
// number of resident requests (each with its own grid size).
//
// One step of cuPentCahnADI (cuPentCahnADI/src/cuPentCahnADI.cu:528-596;
// serial form cuPentSpeedUp/serialCahnADITiming/serialCahnADI.c:1010-1047)
// in four phases:
//   rhs   : cBar = 2c - cOld, Y = (L5x5(cBar) + ((-2/3)(c - cOld) + L3x3(c^3 - c)))^T
//           (the right-hand side is written transposed, through LDS)
//   xsolve: cyclic pentadiagonal solve of every column of Y (pentSolve +
//           solveEnd), X = Y^T with solveFull applied on the way
//   ysolve: cyclic solve of every column of X (pentSolve + solveEnd)
//   final : c_new = cBar + (X - (inv1 x~2 + inv2 x~1)), written to the next
//           field buffer; inside a run of steps it is fused with the next
//           step's rhs (finalRhs)
// Every solve is a thread-per-column sweep (coalesced across the wave) with a
// block of row loads in flight ahead of the recurrence; the transposed store
// of the right-hand side and the transpose after the x solve replace the
// reference's cublasDgeam calls (cuPentCahnADI.cu:552, 566).
// Compiled with -ffp-contract=off: every operation is rounded as written, the
// stencil sums are explicit fma chains (the cuSten order).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ch_kernels.h"
#include "stencil_tile.h"  // xcdRemap

#pragma clang fp contract(off)

namespace custen {
namespace chdev {

// Global-address-space views of the descriptors' field pointers: accesses
// through them are global (not flat) instructions, which the wait counters
// track separately from LDS traffic.
using gdouble = __attribute__((address_space(1))) double;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int findRequest(const CHDesc* d, int count, int item, int which)
{
    // largest r with base(r) <= item; bases ascend with r
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b = which == 0 ? d[mid].tileBase : (which == 2 ? d[mid].colBase : d[mid].trBase);
        if (b <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// t modulo the number of field buffers (2 or 3), t >= 0
__device__ __forceinline__ int ringIdx(int t, int ring) { return ring == 2 ? (t & 1) : (t % 3); }

__device__ __forceinline__ int wrapN(int i, int n)
{
    if (i < 0) return i + n;
    if (i >= n) return i - n;
    return i;
}

// ---------------------------------------------------------------------------
// rhs
// ---------------------------------------------------------------------------
constexpr int RX = kCHTileX, RY = kCHTileY;
constexpr int CBW = RX + 4, CBH = RY + 4;  // cBar tile with a 2-point halo
constexpr int QW = RX + 2, QH = RY + 2;    // u^3 - u tile with a 1-point halo

// The right-hand side of a 64 x 16 tile from the staged cBar (cb, 2-point
// halo) and u^3 - u (q, 1-point halo) and this thread's four values of
// c - cOld (dcs, rows wave*4 .. wave*4+3 of column x0 + lane), written
// transposed to Y.  The caller has synchronised after staging.  The result
// goes through cb's storage, reused once every thread has read it.
static_assert(RX * (RY + 1) <= CBH * CBW, "result tile must fit in the cBar tile");
__device__ __forceinline__ void rhsTile(const CHDesc& D, double* cb, const double* q,
                                        const double (&dcs)[4], int x0, int y0,
                                        gdouble* __restrict__ Yo)
{
    double (*ot)[RY + 1] = reinterpret_cast<double (*)[RY + 1]>(cb);
    const int n = D.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double k23 = -(2.0 / 3.0);
    // The four rows of this lane's column are summed together, weight by
    // weight (each row's chain keeps the cuSten order: j outer, i inner), so
    // each weight is fetched once for the four rows.  Points past the grid
    // are computed from in-bounds LDS values and never stored.
    const int ly0 = wave * 4;
    double lin[4] = {0.0, 0.0, 0.0, 0.0}, nl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const double w = D.wlin[j * 5 + i];
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                lin[rr] = __builtin_fma(w, cb[(ly0 + rr + j) * CBW + lane + i], lin[rr]);
        }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double w = D.wnl[j * 3 + i];
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                nl[rr] = __builtin_fma(w, q[(ly0 + rr + j) * QW + lane + i], nl[rr]);
        }
    double res[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const double t2 = k23 * dcs[rr];
        const double t3 = t2 + nl[rr];
        res[rr] = lin[rr] + t3;
    }
    __syncthreads();  // every thread has read cb
#pragma unroll
    for (int rr = 0; rr < 4; rr++) ot[lane][ly0 + rr] = res[rr];
    __syncthreads();
    // the tile goes out transposed (Y = X^T), 16 consecutive y per x: each
    // store instruction of a wave writes four whole 128-byte segments
#pragma unroll
    for (int i = 0; i < RX * RY / 256; i++) {
        const int idx = i * 256 + (int)threadIdx.x;
        const int xl = idx / RY, yl = idx % RY;
        if (x0 + xl < n && y0 + yl < n) Yo[(int64_t)(x0 + xl) * n + y0 + yl] = ot[xl][yl];
    }
}

__global__ __launch_bounds__(256) void rhsKernel(const CHDesc* __restrict__ descs, int count,
                                                 int step)
{
    __shared__ double cb[CBH * CBW];  // then the tile's result, [x][y]
    __shared__ double q[QH * QW];
    // neighbouring tiles on one XCD, so halo lines are re-read from its L2
    const int item = dev::xcdRemap(blockIdx.x, gridDim.x);
    const int r = findRequest(descs, count, item, 0);
    const CHDesc& D = descs[r];
    const int n = D.n;
    const int tl = step - D.admitStep, ring = D.ring;
    const gdouble* __restrict__ c = (const gdouble*)D.f[ringIdx(tl, ring)];
    const gdouble* __restrict__ co = (const gdouble*)D.f[ringIdx(tl + ring - 1, ring)];
    gdouble* __restrict__ Yo = (gdouble*)D.Y;
    const int local = item - D.tileBase;
    const int tilesX = (n + RX - 1) / RX;
    const int x0 = (local % tilesX) * RX;
    const int y0 = (local / tilesX) * RY;

    // Staging: every thread issues all of its loads before it uses any (a
    // thread's share past the tile repeats the tile's last point, so the
    // loads are unconditional and all in flight together).
    constexpr int kStage = (CBH * CBW + 255) / 256;
    double cv[kStage], ov[kStage];
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int idx = min((int)threadIdx.x + i * 256, CBH * CBW - 1);
        const int rr = idx / CBW, cc = idx - rr * CBW;
        // the points an output reads lie in [-2, n + 1], where one conditional
        // add or subtract wraps (n >= 8); staged points past n + 1 (a tile
        // wider or taller than what is left of the grid) are never read and
        // are clamped to n + 1
        const int gy = wrapN(min(y0 - 2 + rr, n + 1), n);
        const int gx = wrapN(min(x0 - 2 + cc, n + 1), n);
        const int64_t k = (int64_t)gy * n + gx;
        cv[i] = c[k];
        ov[i] = co[k];
    }
    // this thread's own points (4 rows of column x0 + lane), for c - cOld
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = x0 + lane;
    double cOwn[4], oOwn[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int64_t k = (int64_t)min(y0 + wave * 4 + rr, n - 1) * n + min(x, n - 1);
        cOwn[rr] = c[k];
        oOwn[rr] = co[k];
    }
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int idx = (int)threadIdx.x + i * 256;
        if (idx >= CBH * CBW) break;
        const int rr = idx / CBW, cc = idx - rr * CBW;
        const double t = 2.0 * cv[i];
        cb[idx] = t - ov[i];
        if (rr >= 1 && rr < CBH - 1 && cc >= 1 && cc < CBW - 1) {
            const double u2 = cv[i] * cv[i];
            const double u3 = u2 * cv[i];
            q[(rr - 1) * QW + (cc - 1)] = u3 - cv[i];
        }
    }
    double dcs[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) dcs[rr] = cOwn[rr] - oOwn[rr];
    __syncthreads();
    rhsTile(D, cb, q, dcs, x0, y0, Yo);
}

// ---------------------------------------------------------------------------
// transpose (64 x 64 tiles through LDS): X = Y'^T where Y' is Y with the
// column solveFull of the x-direction solves applied (serialCahnADI.c:661-678):
// y_j -= inv1[j] y~2 + inv2[j] y~1, j < n-2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void transposeKernel(const CHDesc* __restrict__ descs, int count)
{
    __shared__ double t[kCHTr][kCHTr + 1];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 3);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int tiles = (n + kCHTr - 1) / kCHTr;
    const int local = item - D.trBase;
    const int R0 = (local / tiles) * kCHTr;  // source rows
    const int C0 = (local % tiles) * kCHTr;  // source columns
    const gdouble* __restrict__ src = (const gdouble*)D.Y;
    gdouble* __restrict__ dst = (gdouble*)D.X;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = C0 + lane;
    const gdouble* inv1 = (const gdouble*)D.fac + 5 * m;
    const gdouble* inv2 = inv1 + m;
    // all loads first (points past the grid repeat its last row or column),
    // then the corrections and the guarded LDS writes
    constexpr int kRowsPerWave = kCHTr / 4;
    const int cc = min(c, n - 1);
    const double e2 = src[(int64_t)(n - 2) * n + cc];
    const double e1 = src[(int64_t)(n - 1) * n + cc];
    double v[kRowsPerWave], i1[kRowsPerWave], i2[kRowsPerWave];
#pragma unroll
    for (int i = 0; i < kRowsPerWave; i++) {
        const int row = R0 + wave + 4 * i;
        v[i] = src[(int64_t)min(row, n - 1) * n + cc];
        i1[i] = inv1[min(row, m - 1)];
        i2[i] = inv2[min(row, m - 1)];
    }
#pragma unroll
    for (int i = 0; i < kRowsPerWave; i++) {
        const int k = wave + 4 * i;
        const int row = R0 + k;
        if (row < n && c < n) {
            double w = v[i];
            if (row < m) w = w - (i1[i] * e2 + i2[i] * e1);
            t[k][lane] = w;
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = wave; k < kCHTr; k += 4) {
        const int row = C0 + k;  // destination row = source column
        const int col = R0 + lane;
        if (row < n && col < n) dst[(int64_t)row * n + col] = t[lane][k];
    }
}

// ---------------------------------------------------------------------------
// column solve: one thread per column (coalesced across the wave).
// pentSolve (serialCahnADI.c:161-241) + solveEnd (:628-655); the column's
// solveFull is applied by the next phase.  which = 0: Y, which = 1: X.
//
// A sweep advances in steps of SR rows.  A wave keeps three steps of its
// column in registers: the step it computes and the two after it, whose
// loads were issued one and two steps earlier; the buffers rotate by
// unrolling three steps, so no register copy waits on a load in flight.  The
// diagonals (shared by every column of a request) of the next step are
// loaded during the current one and written to a two-slot LDS ring at its
// end; rows are read from the ring as broadcasts.  Inside the unrolled loop
// every load and store is unconditional (row indices past the sweep are
// clamped to a valid row and their values unused), so the wait counts the
// compiler computes are exact and no wait drains the loads still in flight.
// Steps that do not fit whole are done one row at a time after the loop.
// The factored rows settle into a constant run (chConstantRuns); the steps
// inside it run as a loop of their own with the run's values as uniform
// constants and no ring, and divide in two operations when the host has
// certified the run's pivot (chDivCertified).
// ---------------------------------------------------------------------------
constexpr int SR = 16;  // rows per step (settleStep takes 16)

// Reciprocal of a diagonal entry for the division below: RN(1/d) when
// 2^-60 <= |d| <= 2^60, else 0 (the row then divides with IEEE division).
__device__ __forceinline__ double divRecip(double d)
{
    const double ad = __builtin_fabs(d);
    return (ad >= 0x1p-60 && ad <= 0x1p60) ? 1.0 / d : 0.0;
}

// n / d, correctly rounded, with y = RN(1/d) precomputed (all columns of a
// request share the matrix, so y is computed once per row, off the
// recurrence's critical path).  q0 = RN(n y) and two corrections
// q <- RN(q + RN(n - d q) y) (each remainder exact by fma) give RN(n / d)
// (Markstein's theorem: the first correction makes q faithful, the second
// rounds it correctly); the dependent chain is five operations against the
// eleven of the scaled IEEE sequence.  The preconditions (no overflow and no
// subnormal intermediate) hold for 2^-900 <= |n| <= 2^900 and the range of d
// in divRecip (y != 0).
__device__ __forceinline__ double divByRecip(double n, double d, double y)
{
    double q = n * y;
    q = __builtin_fma(__builtin_fma(-d, q, n), y, q);
    return __builtin_fma(__builtin_fma(-d, q, n), y, q);
}

// One wait for a whole step of rows: the empty statement takes all SR
// values as operands, so the compiler waits once for the loads that fill
// them (issued two steps earlier, so the wait is normally already satisfied)
// instead of once before each row's first use.
__device__ __forceinline__ void settleStep(const double (&b)[16])
{
    asm volatile("" ::"v"(b[0]), "v"(b[1]), "v"(b[2]), "v"(b[3]), "v"(b[4]), "v"(b[5]), "v"(b[6]),
                 "v"(b[7]), "v"(b[8]), "v"(b[9]), "v"(b[10]), "v"(b[11]), "v"(b[12]), "v"(b[13]),
                 "v"(b[14]), "v"(b[15]));
}

__global__ __launch_bounds__(64, 2) void colSolveKernel(const CHDesc* __restrict__ descs, int count,
                                                     int which)
{
    // ring of diagonals: [slot][array][row of the step]; every lane writes
    // its own row of a slot (rows SR..63 are not read)
    __shared__ double sd[2][4][kCHColBlock];
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 2);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int lane = (int)threadIdx.x;
    // Lanes past the last column repeat the last column: they read the same
    // values and store the same bits, which keeps every access unconditional.
    const int colRaw = (item - D.colBase) * kCHColBlock + lane;
    const bool live = colRaw < n;
    const int col = live ? colRaw : n - 1;
    gdouble* __restrict__ base = (gdouble*)(which ? D.X : D.Y);
    const int64_t s = n;
    const gdouble* __restrict__ ds = (const gdouble*)D.fac;
    const gdouble* __restrict__ dl = ds + m;
    const gdouble* __restrict__ dd = dl + m;
    const gdouble* __restrict__ du = dd + m;
    const gdouble* __restrict__ dw = du + m;
    auto at = [&](int row) -> gdouble& { return base[(int64_t)row * s + col]; };
    // Inside a step every row is addressed as a buffer whose base is the
    // step's first row (the same for all lanes) plus this lane's byte offset
    // of row k of the step, so a load or store needs no address arithmetic
    // of its own.
    char* const gbase = (char*)(which ? D.X : D.Y);
    const size_t stride = (size_t)n * sizeof(double);
    uint32_t off[SR];
#pragma unroll
    for (int k = 0; k < SR; k++) off[k] = (uint32_t)col * 8u + (uint32_t)(k * stride);
    auto rsrc = [](char* p) { return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, 0x7fffffff, 0x00020000); };
    auto load = [&](__amdgpu_buffer_rsrc_t rs, int k) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off[k], 0, 0));
    };
    auto store = [&](__amdgpu_buffer_rsrc_t rs, int k, double v) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, off[k], 0, 0);
    };
    const int nFull = (m - 2) / SR;  // whole steps of each sweep
    // The factors converge: from row cfw on (forward: ds, dl, d) and from row
    // cbw to m-3 (backward: du, dw) every row's are bitwise the same.  Steps
    // inside those runs take them as uniform constants and do without the
    // ring; they run as a loop of their own after (forward) or before
    // (backward) the steps that read the ring.
    const int cfw = D.constFwd1 > 0 ? D.constFwd1 - 1 : m;
    const int cbw = D.constBwd1 > 0 ? D.constBwd1 - 1 : m;
    const int kf = min(cfw, m - 1), kb = min(cbw, m - 3);
    const double k0 = ds[kf], k1 = dl[kf], k2 = dd[kf], ky = divRecip(k2);
    const double ku = du[kb], kw = dw[kb];

    // ---- forward substitution: rows 0 and 1, then rows 2 .. m-1 in steps
    double b2 = at(0) / dd[0];
    at(0) = b2;
    double b1 = (at(1) - dl[1] * b2) / dd[1];
    at(1) = b1;
    if (nFull > 0) {
        // rows 2 + j SR + k of step j (steps past the sweep: the last one)
        const size_t stepBytes = SR * stride, lastRow = (size_t)(2 + (nFull - 1) * SR) * stride;
        auto rows = [&](int j) {
            const size_t o = 2 * stride + (size_t)j * stepBytes;
            return rsrc(gbase + (o < lastRow ? o : lastRow));
        };
        auto loadStep = [&](double (&buf)[SR], __amdgpu_buffer_rsrc_t& p, int j) {
            p = rows(j);
#pragma unroll
            for (int k = 0; k < SR; k++) buf[k] = load(p, k);
        };
        // every row of the staged step has a reciprocal (uniform)
        bool recips = true;
        auto stageDiag = [&](int slot, double p0, double p1, double p2) {
            const double y = divRecip(p2);
            sd[slot][0][lane] = p0;
            sd[slot][1][lane] = p1;
            sd[slot][2][lane] = p2;
            sd[slot][3][lane] = y;
            recips = __all(y != 0.0);
        };
        auto diagRow = [&](int j) { return min(2 + j * SR + min(lane, SR - 1), m - 1); };
        // step j: diagonals of step j+1 and the column rows of step j+2 are
        // loaded, step j is computed from `cur` and ring slot j & 1
        auto step = [&](int j, const double (&cur)[SR], double (&ahead)[SR],
                        const __amdgpu_buffer_rsrc_t& out, __amdgpu_buffer_rsrc_t& outAhead) {
            const int rd = diagRow(j + 1);
            const double p0 = ds[rd], p1 = dl[rd], p2 = dd[rd];
            loadStep(ahead, outAhead, j + 2);
            settleStep(cur);
            const int slot = j & 1;
            double nmin = 0x1p900, nmax = 0x1p-900;
            const double b2s = b2, b1s = b1;
#pragma unroll
            for (int k = 0; k < SR; k++) {
                const double t = cur[k] - sd[slot][0][k] * b2;
                const double num = t - sd[slot][1][k] * b1;
                const double v = divByRecip(num, sd[slot][2][k], sd[slot][3][k]);
                nmin = __builtin_fmin(nmin, __builtin_fabs(num));
                nmax = __builtin_fmax(nmax, __builtin_fabs(num));
                store(out, k, v);
                b2 = b1;
                b1 = v;
            }
            // a numerator outside divByRecip's range (zero, whose sign the
            // corrections would lose, subnormal, huge, Inf) or a row without
            // a reciprocal: this lane repeats the step with IEEE division
            if (__builtin_expect(!(nmin >= 0x1p-900 && nmax <= 0x1p900 && recips), 0)) {
                b2 = b2s;
                b1 = b1s;
                for (int k = 0; k < SR; k++) {
                    const double t = cur[k] - sd[slot][0][k] * b2;
                    const double v = (t - sd[slot][1][k] * b1) / sd[slot][2][k];
                    store(out, k, v);
                    b2 = b1;
                    b1 = v;
                }
            }
            stageDiag(slot ^ 1, p0, p1, p2);
        };
        // a step inside the constant run; `fast`: the pivot is certified for
        // the two-operation division q = fma(x, ky, RN(x kz)) = RN(x / k2)
        // (2^-900 <= |x| <= 2^900), kz the low part of 1/k2
        double kz = 0.0;
        auto stepK = [&](auto fast, int j, const double (&cur)[SR], double (&ahead)[SR],
                        const __amdgpu_buffer_rsrc_t& out, __amdgpu_buffer_rsrc_t& outAhead) {
            loadStep(ahead, outAhead, j + 2);
            settleStep(cur);
            double nmin = 0x1p900, nmax = 0x1p-900;
            const double b2s = b2, b1s = b1;
#pragma unroll
            for (int k = 0; k < SR; k++) {
                const double t = cur[k] - k0 * b2;
                const double num = t - k1 * b1;
                double v;
                if constexpr (decltype(fast)::value) v = __builtin_fma(num, ky, num * kz);
                else v = divByRecip(num, k2, ky);
                nmin = __builtin_fmin(nmin, __builtin_fabs(num));
                nmax = __builtin_fmax(nmax, __builtin_fabs(num));
                store(out, k, v);
                b2 = b1;
                b1 = v;
            }
            if (__builtin_expect(!(nmin >= 0x1p-900 && nmax <= 0x1p900 && ky != 0.0), 0)) {
                b2 = b2s;
                b1 = b1s;
                for (int k = 0; k < SR; k++) {
                    const double t = cur[k] - k0 * b2;
                    const double v = (t - k1 * b1) / k2;
                    store(out, k, v);
                    b2 = b1;
                    b1 = v;
                }
            }
        };
        // steps [0, jc) read the ring, steps [jc, nFull) start at or after cfw
        const int jc = min((cfw - 2 + SR - 1) / SR, nFull);
        double bufA[SR], bufB[SR], bufC[SR];
        __amdgpu_buffer_rsrc_t rA, rB, rC;  // the rows of the steps in bufA, bufB, bufC
        if (jc > 0) {
            {
                const int r0 = diagRow(0);
                stageDiag(0, ds[r0], dl[r0], dd[r0]);
            }
            loadStep(bufA, rA, 0);
            loadStep(bufB, rB, 1);
            int j = 0;
            for (; j + 3 <= jc; j += 3) {
                step(j, bufA, bufC, rA, rC);
                step(j + 1, bufB, bufA, rB, rA);
                step(j + 2, bufC, bufB, rC, rB);
            }
            if (j < jc) step(j++, bufA, bufC, rA, rC);
            if (j < jc) step(j++, bufB, bufA, rB, rA);
        }
        auto constSteps = [&](auto fast) {
            loadStep(bufA, rA, jc);
            loadStep(bufB, rB, jc + 1);
            int j = jc;
            for (; j + 3 <= nFull; j += 3) {
                stepK(fast, j, bufA, bufC, rA, rC);
                stepK(fast, j + 1, bufB, bufA, rB, rA);
                stepK(fast, j + 2, bufC, bufB, rC, rB);
            }
            if (j < nFull) stepK(fast, j++, bufA, bufC, rA, rC);
            if (j < nFull) stepK(fast, j++, bufB, bufA, rB, rA);
        };
        if (jc < nFull) {
            if (D.fastDiv) {
                kz = __builtin_fma(-k2, ky, 1.0) / k2;
                constSteps(std::true_type{});
            } else {
                constSteps(std::false_type{});
            }
        }
    }
    // rows left after the whole steps, one at a time
    for (int i = 2 + nFull * SR; i < m; i++) {
        const double t = at(i) - ds[i] * b2;
        const double v = (t - dl[i] * b1) / dd[i];
        at(i) = v;
        b2 = b1;
        b1 = v;
    }

    // ---- backward substitution: b1 = x[m-1], b2 = b[m-2]; then rows m-3
    // down to 0 in steps
    const double xm1 = b1;
    double a1 = b1;
    double a0 = b2 - du[m - 2] * a1;
    at(m - 2) = a0;
    const double xm2 = a0;
    if (nFull > 0) {
        // step j holds rows m-3 - j SR - k; row k of the step sits at offset
        // SR-1-k above the step's lowest row m-3 - j SR - (SR-1)
        // (offsets counted down from the first step's lowest row, clamped at
        // the last step's)
        const size_t stepBytes = SR * stride;
        const size_t firstRow = (size_t)(m - 3 - (SR - 1)) * stride;
        const size_t lastBack = (size_t)(nFull - 1) * stepBytes;
        auto rows = [&](int j) {
            const size_t b = (size_t)j * stepBytes;
            return rsrc(gbase + firstRow - (b < lastBack ? b : lastBack));
        };
        auto loadStep = [&](double (&buf)[SR], __amdgpu_buffer_rsrc_t& p, int j) {
            p = rows(j);
#pragma unroll
            for (int k = 0; k < SR; k++) buf[k] = load(p, SR - 1 - k);
        };
        auto stageDiag = [&](int slot, double p0, double p1) {
            sd[slot][0][lane] = p0;
            sd[slot][1][lane] = p1;
        };
        auto diagRow = [&](int j) { return max(m - 3 - j * SR - min(lane, SR - 1), 0); };
        // steps [0, jcb) lie in the constant run (their lowest row is at or
        // above cbw), steps [jcb, nFull) read the ring
        const int lowest0 = m - 3 - (SR - 1);
        const int jcb = lowest0 >= cbw ? min((lowest0 - cbw) / SR + 1, nFull) : 0;
        auto step = [&](int j, const double (&cur)[SR], double (&ahead)[SR],
                        const __amdgpu_buffer_rsrc_t& out, __amdgpu_buffer_rsrc_t& outAhead) {
            const int rd = diagRow(j + 1);
            const double p0 = du[rd], p1 = dw[rd];
            loadStep(ahead, outAhead, j + 2);
            settleStep(cur);
            const int slot = (j - jcb) & 1;
#pragma unroll
            for (int k = 0; k < SR; k++) {
                const double t = cur[k] - sd[slot][0][k] * a0;
                const double v = t - sd[slot][1][k] * a1;
                store(out, SR - 1 - k, v);
                a1 = a0;
                a0 = v;
            }
            stageDiag(slot ^ 1, p0, p1);
        };
        auto stepK = [&](int j, const double (&cur)[SR], double (&ahead)[SR],
                        const __amdgpu_buffer_rsrc_t& out, __amdgpu_buffer_rsrc_t& outAhead) {
            loadStep(ahead, outAhead, j + 2);
            settleStep(cur);
#pragma unroll
            for (int k = 0; k < SR; k++) {
                const double t = cur[k] - ku * a0;
                const double v = t - kw * a1;
                store(out, SR - 1 - k, v);
                a1 = a0;
                a0 = v;
            }
        };
        double bufA[SR], bufB[SR], bufC[SR];
        __amdgpu_buffer_rsrc_t rA, rB, rC;  // the rows of the steps in bufA, bufB, bufC
        if (jcb > 0) {
            loadStep(bufA, rA, 0);
            loadStep(bufB, rB, 1);
            int j = 0;
            for (; j + 3 <= jcb; j += 3) {
                stepK(j, bufA, bufC, rA, rC);
                stepK(j + 1, bufB, bufA, rB, rA);
                stepK(j + 2, bufC, bufB, rC, rB);
            }
            if (j < jcb) stepK(j++, bufA, bufC, rA, rC);
            if (j < jcb) stepK(j++, bufB, bufA, rB, rA);
        }
        if (jcb < nFull) {
            {
                const int r0 = diagRow(jcb);
                stageDiag(0, du[r0], dw[r0]);
            }
            loadStep(bufA, rA, jcb);
            loadStep(bufB, rB, jcb + 1);
            int j = jcb;
            for (; j + 3 <= nFull; j += 3) {
                step(j, bufA, bufC, rA, rC);
                step(j + 1, bufB, bufA, rB, rA);
                step(j + 2, bufC, bufB, rC, rB);
            }
            if (j < nFull) step(j++, bufA, bufC, rA, rC);
            if (j < nFull) step(j++, bufB, bufA, rB, rA);
        }
    }
    for (int i = m - 3 - nFull * SR; i >= 0; i--) {
        const double t = at(i) - du[i] * a0;
        const double v = t - dw[i] * a1;
        at(i) = v;
        a1 = a0;
        a0 = v;
    }
    if (!live) return;
    // solveEnd: a0 = x[0], a1 = x[1]
    gdouble* __restrict__ Xc = base + col;
    const double e2 = Xc[(n - 2) * s] - (D.e * a0 + D.a * xm2 + D.b * xm1);
    const double e1 = Xc[(n - 1) * s] - (D.d * a0 + D.e * a1 + D.a * xm1);
    Xc[(n - 2) * s] = D.omega[0] * e2 + D.omega[1] * e1;
    Xc[(n - 1) * s] = D.omega[2] * e2 + D.omega[3] * e1;
}

// ---------------------------------------------------------------------------
// final: column solveFull + findNew, written to the next field buffer (with
// two buffers: over the previous field, each point by the thread that read it)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void finalKernel(const CHDesc* __restrict__ descs, int count,
                                                   int step)
{
    const int item = blockIdx.x;
    const int r = findRequest(descs, count, item, 0);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int tl = step - D.admitStep, ring = D.ring;
    const gdouble* __restrict__ c = (const gdouble*)D.f[ringIdx(tl, ring)];
    // co and out are the same buffer when ring == 2
    const gdouble* co = (const gdouble*)D.f[ringIdx(tl + ring - 1, ring)];
    gdouble* out = (gdouble*)D.f[ringIdx(tl + 1, ring)];
    const gdouble* __restrict__ X = (const gdouble*)D.X;
    const gdouble* inv1 = (const gdouble*)D.fac + 5 * m;
    const gdouble* inv2 = inv1 + m;
    const int local = item - D.tileBase;
    const int tilesX = (n + RX - 1) / RX;
    const int x0 = (local % tilesX) * RX;
    const int y0 = (local / tilesX) * RY;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = x0 + lane;
    if (x >= n) return;
    const double e2 = X[(int64_t)(n - 2) * n + x];
    const double e1 = X[(int64_t)(n - 1) * n + x];
    // all loads first (rows past the grid repeat its last row), then the
    // guarded stores
    double xv[4], cv[4], ov[4], i1[4], i2[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int y = min(y0 + wave * 4 + rr, n - 1);
        const int64_t k = (int64_t)y * n + x;
        xv[rr] = X[k];
        cv[rr] = c[k];
        ov[rr] = co[k];
        const int yi = min(y, m - 1);
        i1[rr] = inv1[yi];
        i2[rr] = inv2[yi];
    }
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int y = y0 + wave * 4 + rr;
        if (y >= n) break;
        double xf = xv[rr];
        if (y < m) xf = xf - (i1[rr] * e2 + i2[rr] * e1);
        const double t = 2.0 * cv[rr];
        const double cBar = t - ov[rr];
        out[(int64_t)y * n + x] = cBar + xf;
    }
}

// ---------------------------------------------------------------------------
// finalRhs: final of step `step` and rhs of step + 1 (three field buffers).
// The new field is pointwise in X, c and cOld, so it is computed over the
// tile and its 2-point halo from those (the halo points as their own tiles
// compute them: the same operations on the same values), written to the
// third buffer at the tile's points, and the next right-hand side is formed
// from it and c without reading it back.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void finalRhsKernel(const CHDesc* __restrict__ descs, int count,
                                                      int step)
{
    __shared__ double cb[CBH * CBW];  // then the tile's result, [x][y]
    __shared__ double q[QH * QW];
    __shared__ double dn[RY][RX];           // c_new - c at the tile's points
    __shared__ double ce2[CBW], ce1[CBW];   // X rows n-2 and n-1 at the staged columns
    __shared__ double ri1[CBH], ri2[CBH];   // inv1, inv2 at the staged rows
    // neighbouring tiles on one XCD, so halo lines are re-read from its L2
    const int item = dev::xcdRemap(blockIdx.x, gridDim.x);
    const int r = findRequest(descs, count, item, 0);
    const CHDesc& D = descs[r];
    const int n = D.n, m = n - 2;
    const int tl = step - D.admitStep, ring = D.ring;
    const gdouble* __restrict__ c = (const gdouble*)D.f[ringIdx(tl, ring)];
    const gdouble* __restrict__ co = (const gdouble*)D.f[ringIdx(tl + ring - 1, ring)];
    gdouble* __restrict__ out = (gdouble*)D.f[ringIdx(tl + 1, ring)];
    const gdouble* __restrict__ X = (const gdouble*)D.X;
    gdouble* __restrict__ Yo = (gdouble*)D.Y;
    const gdouble* inv1 = (const gdouble*)D.fac + 5 * m;
    const gdouble* inv2 = inv1 + m;
    const int local = item - D.tileBase;
    const int tilesX = (n + RX - 1) / RX;
    const int x0 = (local % tilesX) * RX;
    const int y0 = (local / tilesX) * RY;
    const int tid = (int)threadIdx.x;

    // staged points as in rhsKernel: [-2, n + 1] wrapped, clamped past n + 1
    constexpr int kStage = (CBH * CBW + 255) / 256;
    double xv[kStage], cv[kStage], ov[kStage];
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int idx = min(tid + i * 256, CBH * CBW - 1);
        const int rr = idx / CBW, cc = idx - rr * CBW;
        const int gy = wrapN(min(y0 - 2 + rr, n + 1), n);
        const int gx = wrapN(min(x0 - 2 + cc, n + 1), n);
        const int64_t k = (int64_t)gy * n + gx;
        xv[i] = X[k];
        cv[i] = c[k];
        ov[i] = co[k];
    }
    if (tid < CBW) {
        const int gx = wrapN(min(x0 - 2 + tid, n + 1), n);
        ce2[tid] = X[(int64_t)(n - 2) * n + gx];
        ce1[tid] = X[(int64_t)(n - 1) * n + gx];
    } else if (tid < CBW + CBH) {
        const int rr = tid - CBW;
        const int yi = min(wrapN(min(y0 - 2 + rr, n + 1), n), m - 1);
        ri1[rr] = inv1[yi];
        ri2[rr] = inv2[yi];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int idx = tid + i * 256;
        if (idx >= CBH * CBW) break;
        const int rr = idx / CBW, cc = idx - rr * CBW;
        const int gy = wrapN(min(y0 - 2 + rr, n + 1), n);
        // final (as finalKernel)
        double xf = xv[i];
        if (gy < m) xf = xf - (ri1[rr] * ce2[cc] + ri2[rr] * ce1[cc]);
        const double t = 2.0 * cv[i];
        const double cBar = t - ov[i];
        const double cn = cBar + xf;
        // staging of the next step's rhs (as rhsKernel, c -> cn, cOld -> c)
        const double t2 = 2.0 * cn;
        cb[idx] = t2 - cv[i];
        if (rr >= 1 && rr < CBH - 1 && cc >= 1 && cc < CBW - 1) {
            const double u2 = cn * cn;
            const double u3 = u2 * cn;
            q[(rr - 1) * QW + (cc - 1)] = u3 - cn;
        }
        if (rr >= 2 && rr < 2 + RY && cc >= 2 && cc < 2 + RX) {
            const int y = y0 + rr - 2, x = x0 + cc - 2;
            if (y < n && x < n) out[(int64_t)y * n + x] = cn;
            dn[rr - 2][cc - 2] = cn - cv[i];
        }
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    double dcs[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) dcs[rr] = dn[wave * 4 + rr][lane];
    rhsTile(D, cb, q, dcs, x0, y0, Yo);
}

}  // namespace chdev

hipError_t chLaunchRhs(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::rhsKernel, dim3(totalTiles), dim3(256), 0, st, descs, count, step);
    return hipGetLastError();
}

hipError_t chLaunchTranspose(const CHDesc* descs, int count, int totalTiles, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::transposeKernel, dim3(totalTiles), dim3(256), 0, st, descs, count);
    return hipGetLastError();
}

hipError_t chLaunchColSolve(const CHDesc* descs, int count, int totalColBlocks, int which,
                            hipStream_t st)
{
    if (count <= 0 || totalColBlocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::colSolveKernel, dim3(totalColBlocks), dim3(kCHColBlock), 0, st,
                       descs, count, which);
    return hipGetLastError();
}

hipError_t chLaunchFinalRhs(const CHDesc* descs, int count, int totalTiles, int step,
                            hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::finalRhsKernel, dim3(totalTiles), dim3(256), 0, st, descs, count,
                       step);
    return hipGetLastError();
}

hipError_t chLaunchFinal(const CHDesc* descs, int count, int totalTiles, int step, hipStream_t st)
{
    if (count <= 0 || totalTiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(chdev::finalKernel, dim3(totalTiles), dim3(256), 0, st, descs, count, step);
    return hipGetLastError();
}

}  // namespace custen
