// stencil_band.h -- the band-walker stencil kernel: one wave streams a band
// of rows down its own 128-column strip, with no workgroup barrier.
//
// Rows go from HBM straight into LDS (global_load_lds: each lane's 16 or 4
// bytes land at LDS base + 16 * lane or 4 * lane, so one instruction stages a
// 128-column FP64 row body and two stage an FP32 one; one 4-byte load per
// side by the lanes that cover HLP columns stages the halo).  The wave keeps
// a ring of S row slots,
//     S = (D + 1) * RPW + V - 1,
// and while it computes chunk c (RPW output rows from RPW + V - 1 input rows)
// the rows of chunks c+1 .. c+D are in flight into the other slots, so no
// registers are spent on staging and the V-1 rows shared by consecutive
// chunks are never copied.  Completion is tracked with the vector-memory
// counter, which retires loads and stores in issue order: before chunk c the
// wave waits until no more than (loads issued after chunk c's) + (stores
// issued after chunk c's) operations are outstanding.  A chunk whose stores
// did not take the one-store-per-row path (write-masked or ragged rows)
// makes the following waits drain the counter instead.
//
// Arithmetic is the fma chain of the other stencil kernels (row j outer,
// column i inner, from 0.0), so every configuration gives identical bits.
// The point-wise transform of the cube-minus point function (u^3 - u,
// unfused) is applied to each value as it is read from the ring; the strip
// kernel applies the same function while staging, so the values agree.
#pragma once

#include "stencil_tile.h"

#pragma clang fp contract(off)

namespace custen {
namespace dev {

// LDS (address space 3) views: the ring is addressed through them from the
// kernel down, so the compiler never has to turn a generic pointer into an
// LDS address.
typedef __attribute__((address_space(3))) void LdsVoid;
template <typename T> using Lds = __attribute__((address_space(3))) T;

// s_waitcnt with vmcnt = N and the other counters unconstrained (gfx9
// encoding: vmcnt[3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
template <int N>
__device__ __forceinline__ void waitVmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt is a 6-bit counter");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
    __asm__ __volatile__("" ::: "memory");
}

// Wait for a chunk given la = chunks of loads and sa = chunks of stores
// issued after it (la < D, sa <= D): the immediate is la*NL + sa*NS.
template <int NL, int NS, int D, int A = 0, int B = 0>
__device__ __forceinline__ void waitChunk(int la, int sa)
{
    if constexpr (A >= D) {
        waitVmcnt<0>();
    } else if constexpr (B > D) {
        waitChunk<NL, NS, D, A + 1, 0>(la, sa);
    } else {
        if (la == A && sa == B) waitVmcnt<A * NL + B * NS>();
        else waitChunk<NL, NS, D, A, B + 1>(la, sa);
    }
}

// Direct-to-LDS loads use the default cache policy.  Non-temporal row loads
// were measured: 6-10 % faster when the same input is applied again and
// again (the unchanged input then stays partly in the last-level cache), but
// within 2 % either way in a time-stepping loop that alternates input and
// output, and 10-40 % slower for tall stencils and FP32 fields.  Since the
// autotuner times repeated applies of one input, that policy would also
// mislead it, so it is not used.
constexpr int kDmaDefault = 0;

template <int POLICY>
__device__ __forceinline__ void ldsDma16(const Glob<double>* g, Lds<double>* l)
{
    __builtin_amdgcn_global_load_lds((const Glob<void>*)g, (LdsVoid*)l, 16, 0, POLICY);
}
template <int POLICY, typename T>
__device__ __forceinline__ void ldsDma4(const Glob<unsigned>* g, Lds<T>* l)
{
    __builtin_amdgcn_global_load_lds((const Glob<void>*)g, (LdsVoid*)l, 4, 0, POLICY);
}

// FP64 stencils of more than 36 points have more weights than the scalar
// registers hold next to the walker's own state; held for the whole walk they
// are spilled and read back lane by lane.  The unmasked (FULL) walk of those
// stencils instead loads each weight row from the scalar cache just before
// its FMAs (the weights are not written during a launch, so the constant
// address space applies).  The masked walk keeps resident weights: with its
// store masks in the loop the compiler spilled more when streaming.
template <typename T, int H, int V, bool FULL>
constexpr bool kStreamWeights = FULL && sizeof(T) == 8 && H * V > 36;

// FP32 stencils of more than 36 points: the weights (converted from the
// FP64 array on the vector unit) stay in vector registers, 49 or 81 of them,
// instead of being moved to scalar registers that then spill.
template <typename T, int H, int V>
constexpr bool kWeightsInVgprs = sizeof(T) == 4 && H * V > 36;

template <int H>
__device__ __forceinline__ void loadWeightRow(const double* weights, int j, double (&wr)[H],
                                              double after)
{
    // The empty asm makes the address opaque at each use and orders it after
    // `after` (the accumulator the row updates), so the loads stay next to
    // the FMAs that consume them instead of being hoisted out of the walk.
    uint64_t p = (uint64_t)weights;
    __asm__ __volatile__("" : "+s"(p) : "v"(after));
    typedef __attribute__((address_space(4))) const double CDouble;
    const CDouble* q = (const CDouble*)p;
#pragma unroll
    for (int i = 0; i < H; i++) wr[i] = q[j * H + i];
}

// Ring geometry of a band-walker instantiation (2 columns per lane).  A row
// body is one 16-byte load per lane for FP64 and two 4-byte loads (columns
// lane and 64 + lane) for FP32; each halo side is one 4-byte load by the
// lanes that cover HLP columns.
template <typename T, int H, int V, int RPW>
struct BandShape {
    static constexpr int D = 8 / RPW;                  // chunks in flight ahead
    static constexpr int HLP = ((H - 1) + 1) & ~1;     // halo columns staged per side
    static constexpr int LD = kTW + 2 * HLP;           // slot length (elements)
    static constexpr int S = (D + 1) * RPW + V - 1;    // slots in the ring
    static constexpr int BODY = sizeof(T) == 8 ? 1 : 2;      // body loads per row
    static constexpr int HALO_LANES = HLP * (int)sizeof(T) / 4;  // dwords per side
    static constexpr int NL = RPW * (BODY + (H > 1 ? 2 : 0));  // load instructions per chunk
    static constexpr int LDS_BYTES = S * LD * (int)sizeof(T);
};

// Band walk of wave g of a launch (one wave per workgroup): interior strip
// stripTx0 + g % nStripTx (or, with edgeList, edgeTx[g % nStripTx]) over
// region rows [band * RPW * chunks, ...) with band = g / nStripTx.  `ring` is the wave's BandShape::S x LD LDS ring and
// `w` the H*V weights; X the point-wise transform of the values read.
// Requirements (checked by the host): the strip is interior (full 128
// columns, pair-aligned rows), and wrapped row indices need at most one
// correction for every row the band touches.  FULL (also decided by the
// host): every point of every walked strip is written and the region height
// is a multiple of RPW, so each chunk stores RPW whole rows.  With
// kStreamWeights the weights come from a.weights and `w` is not read.
template <typename T, int H, int V, int RPW, bool FULL, Xform X = Xform::Identity>
__device__ __forceinline__ void bandWalk(const StencilLaunch& a, const T (&w)[H * V],
                                         Lds<T>* __restrict__ ring, int g)
{
    using BS = BandShape<T, H, V, RPW>;
    constexpr int D = BS::D, HLP = BS::HLP, LD = BS::LD, S = BS::S, NL = BS::NL;
    constexpr int HL = BS::HALO_LANES;

    const int lane = threadIdx.x;
    const int tx = a.edgeList ? a.edgeTx[g % a.nStripTx] : a.stripTx0 + g % a.nStripTx;
    const int band = g / a.nStripTx;
    const int bandRows = RPW * a.chunks;
    const int yS = band * bandRows;
    if (yS >= a.ny) return;
    const int yE = min(yS + bandRows, a.ny);
    const int nChunks = (yE - yS + RPW - 1) / RPW;

    const Glob<T>* __restrict__ in = globalPtr<T>(a.in);
    Glob<T>* __restrict__ out = globalPtr<T>(a.out);
    const int nx = a.nx;
    const int x0 = tx * kTW;
    const int L = a.L;
    const int R = H - 1 - L;
    const bool fullRow = x0 >= a.wx0 && x0 + kTW <= a.wx1;

    // halo columns staged for this lane (lanes < HL, one dword each)
    const int hcol = sizeof(T) == 8 ? lane >> 1 : lane;
    const int hword = sizeof(T) == 8 ? lane & 1 : 0;
    int colL = x0 - HLP + hcol;
    int colR = x0 + kTW + hcol;
    if (a.xWrap) {
        if (colL < 0) colL += nx;
        if (colR >= nx) colR -= nx;
    } else {
        colL = colL < 0 ? 0 : colL;
        colR = colR >= nx ? nx - 1 : colR;
    }

    // Rows are loaded in order; syNext is the source row of the next one.
    const RowMap rows = a.rows;
    int gyNext = yS - a.T;
    int syNext = mapRowNear(rows, gyNext);
    auto loadRow = [&](int slot) {
        const Glob<T>* row = in + (int64_t)syNext * a.ldIn;
        gyNext++;
        if (rows.wrap) syNext = syNext + 1 == rows.rowWrap ? 0 : syNext + 1;
        else syNext = min(max(gyNext, rows.rowMin), rows.rowMax);
        Lds<T>* l = ring + slot * LD;
        if constexpr (sizeof(T) == 8) {
            ldsDma16<kDmaDefault>((const Glob<double>*)(row + x0 + 2 * lane),
                                  (Lds<double>*)(l + HLP));
        } else {
            ldsDma4<kDmaDefault>((const Glob<unsigned>*)(row + x0 + lane), l + HLP);
            ldsDma4<kDmaDefault>((const Glob<unsigned>*)(row + x0 + kLanes + lane),
                                 l + HLP + kLanes);
        }
        if constexpr (H > 1) {
            if (lane < HL) {
                ldsDma4<kDmaDefault>((const Glob<unsigned>*)(row + colL) + hword, l);
                ldsDma4<kDmaDefault>((const Glob<unsigned>*)(row + colR) + hword, l + HLP + kTW);
            }
        }
    };
    auto loadChunk = [&](int slot0) {
#pragma unroll
        for (int r = 0; r < RPW; r++) {
            int sl = slot0 + r;
            if (sl >= S) sl -= S;
            loadRow(sl);
        }
    };

    // Chunk c computes region rows [yS + c*RPW, +RPW) from relative input rows
    // [c*RPW, c*RPW + RPW + V - 1) (relative row 0 is yS - T), held in slots
    // (relative row) mod S.  Its loads are the rows [c*RPW + V-1, (c+1)*RPW + V-1);
    // chunk 0 also loads rows [0, V-1) first.
#pragma unroll
    for (int r = 0; r < V - 1; r++) loadRow(r);
    int slotLoad = V - 1;
    for (int c = 0; c < D; c++) {
        if (c < nChunks) loadChunk(slotLoad);
        slotLoad += RPW;
        if (slotLoad >= S) slotLoad -= S;
    }

    int slotBase = 0;
    unsigned inexact = 0;  // bit i: chunk c-1-i did not store one row per instruction
    for (int c = 0; c < nChunks; c++) {
        if (!FULL && (inexact & ((1u << D) - 1u))) {
            waitVmcnt<0>();
        } else {
            const int la = min(c + D - 1, nChunks - 1) - c;
            const int sa = min(c, D);
            waitChunk<NL, RPW, D>(la, sa);
        }
        // chunk c + D reuses the slots of chunk c - 1's first RPW rows
        if (c + D < nChunks) loadChunk(slotLoad);
        slotLoad += RPW;
        if (slotLoad >= S) slotLoad -= S;

        T acc[RPW][2];
#pragma unroll
        for (int r = 0; r < RPW; r++) acc[r][0] = acc[r][1] = T(0);
#pragma unroll
        for (int ir = 0; ir < RPW + V - 1; ir++) {
            int sl = slotBase + ir;
            if (sl >= S) sl -= S;
            const Lds<T>* row = ring + sl * LD + (HLP - L) + 2 * lane;
            T v[H + 1];
            // Pair reads only for stencils of at most 25 points: the wider
            // ones were measured slower with a second read path in the loop.
            if (H * V <= 25 && ((HLP - L) & 1) == 0 && (H + 1) % 2 == 0) {
                // window pair-aligned: one 2-element read per pair (FP64:
                // ds_read_b128, 16 lanes x 16 B cover all banks once)
                using V2 = typename Vec2<T>::type;
                const Lds<V2>* rp = (const Lds<V2>*)row;
#pragma unroll
                for (int k = 0; k < (H + 1) / 2; k++) {
                    const V2 p = rp[k];
                    v[2 * k] = xformValue<T, X>(p.x);
                    v[2 * k + 1] = xformValue<T, X>(p.y);
                }
            } else {
#pragma unroll
                for (int k = 0; k < H + 1; k++) v[k] = xformValue<T, X>(row[k]);
            }
#pragma unroll
            for (int r = 0; r < RPW; r++) {
                const int j = ir - r;
                if (j >= 0 && j < V) {
                    T wr[H];
                    if constexpr (kStreamWeights<T, H, V, FULL>) {
                        loadWeightRow<H>(a.weights, j, wr, acc[r][1]);
                    } else {
#pragma unroll
                        for (int i = 0; i < H; i++) wr[i] = w[j * H + i];
                    }
#pragma unroll
                    for (int i = 0; i < H; i++) {
                        acc[r][0] = fmaT(wr[i], v[i], acc[r][0]);
                        acc[r][1] = fmaT(wr[i], v[i + 1], acc[r][1]);
                    }
                }
            }
        }

        if constexpr (FULL) {
#pragma unroll
            for (int r = 0; r < RPW; r++) {
                using V2 = typename Vec2<T>::type;
                const V2 o = {acc[r][0], acc[r][1]};
                __builtin_nontemporal_store(
                    o, reinterpret_cast<Glob<V2>*>(out + (int64_t)(yS + c * RPW + r) * a.ldOut +
                                                   x0 + 2 * lane));
            }
        } else {
            bool exact = fullRow;
#pragma unroll
            for (int r = 0; r < RPW; r++) {
                const int gy = yS + c * RPW + r;
                const int gyg = a.yGlobal0 + gy;
                if (gy >= yE || gyg < a.wy0 || gyg >= a.wy1) {
                    exact = false;
                    continue;
                }
                storeLane<T, 2>(out + (int64_t)gy * a.ldOut, x0 + 2 * lane, acc[r], fullRow, nx,
                                a.wx0, a.wx1, a.zeroRight, R);
            }
            inexact = (inexact << 1) | (exact ? 0u : 1u);
        }
        slotBase += RPW;
        if (slotBase >= S) slotBase -= S;
    }
}

// Single-region launch: one wave per workgroup, waves in XCD-aware order.
template <typename T, int H, int V, int RPW, bool FULL, Xform X = Xform::Identity>
__global__ __launch_bounds__(kLanes) void stencilBand(StencilLaunch a)
{
    using BS = BandShape<T, H, V, RPW>;
    __shared__ __attribute__((aligned(16))) T ring[BS::S * BS::LD];
    T w[H * V];
    if constexpr (!kStreamWeights<T, H, V, FULL>) {
#pragma unroll
        for (int k = 0; k < H * V; k++) w[k] = static_cast<T>(a.weights[k]);
    }
    if constexpr (kWeightsInVgprs<T, H, V>) {
        // an empty asm with a vector-register result: the compiler treats
        // the value as per-lane and keeps it there
#pragma unroll
        for (int k = 0; k < H * V; k++) {
            T o;
            __asm__ __volatile__("" : "=v"(o) : "0"(w[k]));
            w[k] = o;
        }
    }
    bandWalk<T, H, V, RPW, FULL, X>(a, w, (Lds<T>*)ring, xcdRemap(blockIdx.x, gridDim.x));
}

}  // namespace dev
}  // namespace custen
