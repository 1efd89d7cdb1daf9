// ch.cpp -- Cahn-Hilliard ADI on the GPU: per-request setup, a batched step
// engine, the single-simulation API (custenCH*) and the request server with
// continuous batching (custenCHServer*).
//
// Setup restates cuPentCahnADI/src/cuPentCahnADI.cu:204-230, 368-446 and
// BatchHyper.cu:24-514 (serial form: serialCahnADI.c:40-417, 776-970):
// sigma = 2 dt D gamma / (3 dx^4), diagonals [s, -4s, 1+6s, -4s, s], the
// factorisation of the leading (n-2) block, inv1 = E^-1 f1, inv2 = E^-1 f2 and
// omega = the inverse of the 2x2 Schur complement.  All n systems of one grid
// share one matrix, so the factored diagonals are stored once per request
// (7 (n-2) doubles) instead of replicated per system as the reference does.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuSten.h"
#include "../../include/custen_ch.h"
#include "ch_kernels.h"
#include "ch_plain.h"
#include "ch_setup.h"
#include "plan_internal.h"

#pragma STDC FP_CONTRACT OFF

namespace custen {
namespace {

using Clock = std::chrono::steady_clock;

struct CHCoeffs {
    int n = 0, m = 0;
    double dx = 0, dt = 0, sl = 0, snl = 0;
    double a = 0, b = 0, c = 0, d = 0, e = 0;
    double omega[4] = {0, 0, 0, 0};
    double wlin[25];
    double wnl[9];
    std::vector<double> fac;  // ds, dl, d, du, dw, inv1, inv2
    int constFwd = 0, constBwd = 0;  // chConstantRuns(fac)
    int fastDiv = 0;                 // chDivCertified(pivot of row constFwd)
};

// Validation and the constants of the scheme (everything but the factors).
bool chConstants(const custenCHParams& p, CHCoeffs& co, std::string& why)
{
    if (p.n < 8 || p.n > kCHMaxN) {
        why = "n must be in [8, " + std::to_string(kCHMaxN) + "]";
        return false;
    }
    if (!(p.lx > 0) || !(p.dtOverDx > 0) || !std::isfinite(p.D) || !std::isfinite(p.gamma)) {
        why = "lx and dtOverDx must be positive, D and gamma finite";
        return false;
    }
    const int n = p.n, m = n - 2;
    co.n = n;
    co.m = m;
    co.dx = p.lx / n;
    co.dt = p.dtOverDx * co.dx;
    const double dx = co.dx, dt = co.dt;
    co.sl = 2.0 * dt * p.D * p.gamma / (3.0 * (dx * dx * dx * dx));
    co.snl = (dt / 3.0) * p.D * (2.0 / (dx * dx));
    const double s = co.sl;
    co.a = s;
    co.b = -4 * s;
    co.c = 1 + 6 * s;
    co.d = -4 * s;
    co.e = s;
    static const double lin[25] = {0, 0, -1, 0, 0,  0, -2, 8, -2, 0,  -1, 8, -20, 8, -1,
                                   0, -2, 8, -2, 0, 0, 0, -1, 0, 0};
    static const double nl[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
    for (int k = 0; k < 25; k++) co.wlin[k] = lin[k] * s;
    for (int k = 0; k < 9; k++) co.wnl[k] = nl[k] * co.snl;
    return true;
}

bool chSetup(const custenCHParams& p, CHCoeffs& co, std::string& why)
{
    if (!chConstants(p, co, why)) return false;
    co.fac.assign((size_t)7 * co.m, 0.0);
    chFactor(co.fac.data(), co.m, co.a, co.b, co.c, co.d, co.e, co.omega);
    chConstantRuns(co.fac.data(), co.m, co.constFwd, co.constBwd);
    co.fastDiv = chDivCertified(co.fac[(size_t)2 * co.m + co.constFwd]) ? 1 : 0;
    return true;
}

// Device buffers of one request, recycled by grid size.
struct CHBuffers {
    int n = 0;
    double* f[3] = {nullptr, nullptr, nullptr};  // field buffers, used in turn
    double* X = nullptr;
    double* Y = nullptr;
    double* fac = nullptr;
};

size_t footprintBytes(int n)
{
    const size_t nn = (size_t)n * n;
    return 5 * nn * sizeof(double) + (size_t)7 * (n - 2) * sizeof(double);
}

class BufferPool {
public:
    explicit BufferPool(size_t keepBytes) : keep_(keepBytes) {}
    ~BufferPool()
    {
        for (auto& kv : free_)
            for (auto& b : kv.second) release(b);
    }
    hipError_t get(int n, CHBuffers& out)
    {
        auto it = free_.find(n);
        if (it != free_.end() && !it->second.empty()) {
            out = it->second.back();
            it->second.pop_back();
            held_ -= footprintBytes(n);
            return hipSuccess;
        }
        out = CHBuffers{};
        out.n = n;
        const size_t nn = (size_t)n * n;
        hipError_t e;
        if ((e = hipMalloc(&out.f[0], nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&out.f[1], nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&out.f[2], nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&out.X, nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&out.Y, nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&out.fac, (size_t)7 * (n - 2) * sizeof(double))) != hipSuccess) {
            release(out);
            return e;
        }
        return hipSuccess;
    }
    void put(const CHBuffers& b)
    {
        if (held_ + footprintBytes(b.n) > keep_) {
            CHBuffers c = b;
            release(c);
            return;
        }
        free_[b.n].push_back(b);
        held_ += footprintBytes(b.n);
    }

private:
    static void release(CHBuffers& b)
    {
        for (double* p : {b.f[0], b.f[1], b.f[2], b.X, b.Y, b.fac})
            if (p) (void)hipFree(p);
        b.f[0] = b.f[1] = b.f[2] = b.X = b.Y = b.fac = nullptr;
    }
    size_t keep_;
    size_t held_ = 0;
    std::map<int, std::vector<CHBuffers>> free_;
};

struct CHRequest {
    int id = -1;
    custenCHParams params{};
    CHCoeffs co;
    int steps = 0;
    int done = 0;
    int admitStep = 0;
    CHBuffers buf;
    std::vector<double> c0;
    std::vector<double> cOld0;  // empty: cOld starts equal to c0
    bool staged = false;        // buf already holds c0, cOld0 and the factors
    std::vector<double> result;
    double* pinnedResult = nullptr;  // the result, when held in pinned memory
    int state = 0;  // 0 queued, 1 resident (or its result in transit), 2 finished
    Clock::time_point tSubmit, tAdmit, tFinish;
};

// Writes a request's fields and factors into its device buffers: c0 into
// f[0] (the current field at the request's first step) and cOld0 into f[2]
// (the previous one), or a copy of c0 when cOld0 is null.
hipError_t loadRequest(const CHRequest* r, const double* c0, const double* cOld0, hipStream_t st)
{
    const size_t bytes = (size_t)r->co.n * r->co.n * sizeof(double);
    hipError_t e;
    if ((e = hipMemcpyAsync(r->buf.f[0], c0, bytes, hipMemcpyDefault, st)) != hipSuccess) return e;
    if (cOld0) e = hipMemcpyAsync(r->buf.f[2], cOld0, bytes, hipMemcpyDefault, st);
    else e = hipMemcpyAsync(r->buf.f[2], r->buf.f[0], bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(r->buf.fac, r->co.fac.data(), r->co.fac.size() * sizeof(double),
                          hipMemcpyHostToDevice, st);
}

// Advances every resident request by one step with one launch per phase.
class CHEngine {
public:
    explicit CHEngine(int device) : device_(device) {}
    ~CHEngine()
    {
        for (BatchGraph& g : graphs_)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        releaseDescs();
    }

    int step() const { return step_; }

    // cOld0 == nullptr: cOld starts equal to c (cuPentCahnADI.cu:295-305);
    // otherwise the pair (c0, cOld0) resumes a simulation exactly.
    hipError_t admit(CHRequest* r, const double* c0, const double* cOld0, hipStream_t st)
    {
        hipError_t e = loadRequest(r, c0, cOld0, st);
        if (e != hipSuccess) return e;
        adopt(r);
        return hipSuccess;
    }

    // A request whose buffers already hold its fields and factors (written
    // by loadRequest, on a stream the host has synchronised with).
    void adopt(CHRequest* r)
    {
        r->admitStep = step_;
        active_.push_back(r);
        dirty_ = true;
    }

    void remove(CHRequest* r)
    {
        active_.erase(std::remove(active_.begin(), active_.end(), r), active_.end());
        dirty_ = true;
    }

    const std::vector<CHRequest*>& active() const { return active_; }

    // Current and previous field of a resident request: the three buffers
    // are used in turn, each step writing the new field to the one after
    // the current.
    double* current(const CHRequest* r) const { return r->buf.f[(step_ - r->admitStep) % 3]; }
    double* previous(const CHRequest* r) const
    {
        return r->buf.f[(step_ - r->admitStep + 2) % 3];
    }

    // A step done outside the engine (the plain path) that kept the field in
    // place: advance the step counter without flipping the request's roles.
    void noteExternalStep(CHRequest* r)
    {
        step_++;
        r->admitStep++;
        dirty_ = true;
    }

    // k steps of all resident requests, launched directly on `st`.
    hipError_t run(hipStream_t st, int k)
    {
        if (active_.empty() || k <= 0) return hipSuccess;
        hipError_t e;
        if (dirty_ && (e = upload(st)) != hipSuccess) return e;
        if ((e = launchSteps(st, step_, k)) != hipSuccess) return e;
        step_ += k;
        return hipSuccess;
    }

    // k steps of all resident requests as one launch: the step kernels of the
    // k steps are captured once into a HIP graph and replayed while the
    // resident set (and so the descriptors) and k stay the same.  The
    // kernels use a step number only modulo 3 relative to each request's
    // admission (which field buffer is current), so a graph serves every
    // batch whose first step has the residue it was captured at.  Replaying
    // one graph costs the host one launch instead of 4 k + 1.  Consecutive batches alternate between two
    // graphs: launching a graph whose previous launch is still running makes
    // the host wait for it.
    hipError_t stepBatch(hipStream_t st, int k)
    {
        if (active_.empty() || k <= 0) return hipSuccess;
        hipError_t e;
        if (dirty_ && (e = upload(st)) != hipSuccess) return e;
        BatchGraph& g = graphs_[launches_ & 1];
        if (!g.exec || g.k != k || g.version != version_ || g.parity != step_ % 3) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return e;
            const hipError_t le = launchSteps(st, step_, k);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(st, &graph);
            if (le != hipSuccess) e = le;
            if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                g.exec = nullptr;
                return e;
            }
            g.k = k;
            g.version = version_;
            g.parity = step_ % 3;
        }
        if ((e = hipGraphLaunch(g.exec, st)) != hipSuccess) return e;
        launches_++;
        step_ += k;
        return hipSuccess;
    }

private:
    void releaseDescs()
    {
        if (dDescs_) (void)hipFree(dDescs_);
        dDescs_ = nullptr;
        for (int i = 0; i < 2; i++) {
            if (staged_[i]) (void)hipEventSynchronize(staged_[i]);
            if (hDescs_[i]) (void)hipHostFree(hDescs_[i]);
            if (staged_[i]) (void)hipEventDestroy(staged_[i]);
            hDescs_[i] = nullptr;
            staged_[i] = nullptr;
            stagedUsed_[i] = false;
        }
    }

    // Steps s0 .. s0 + k - 1: the rhs of the first, then per step the two
    // solves and the transpose, and the final phase fused with the next
    // step's rhs (the last step's final alone).
    hipError_t launchSteps(hipStream_t st, int s0, int k)
    {
        hipError_t e;
        if ((e = chLaunchRhs(dDescs_, count_, tiles_, s0, st)) != hipSuccess) return e;
        for (int i = 0; i < k; i++) {
            const int s = s0 + i;
            if ((e = chLaunchColSolve(dDescs_, count_, colBlocks_, 0, st)) != hipSuccess) return e;
            if ((e = chLaunchTranspose(dDescs_, count_, trTiles_, st)) != hipSuccess) return e;
            if ((e = chLaunchColSolve(dDescs_, count_, colBlocks_, 1, st)) != hipSuccess) return e;
            e = i + 1 < k ? chLaunchFinalRhs(dDescs_, count_, tiles_, s, st)
                          : chLaunchFinal(dDescs_, count_, tiles_, s, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

    hipError_t upload(hipStream_t st)
    {
        const int n = (int)active_.size();
        hipError_t e;
        if (n > cap_) {
            // the previous descriptor upload may still be read by queued launches
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            releaseDescs();
            cap_ = std::max(n, 2 * cap_);
            if ((e = hipMalloc(&dDescs_, sizeof(CHDesc) * cap_)) != hipSuccess) return e;
            for (int i = 0; i < 2; i++) {
                if ((e = hipHostMalloc(&hDescs_[i], sizeof(CHDesc) * cap_, 0)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&staged_[i], hipEventDisableTiming)) != hipSuccess)
                    return e;
                stagedUsed_[i] = false;
            }
        }
        // Two pinned staging copies used in turn; one is rewritten only after
        // the copy that last read it has run.  The device array needs no such
        // wait: the copy into it is queued behind every launch that reads it.
        const int u = (int)(uploads_++ & 1);
        if (stagedUsed_[u] && (e = hipEventSynchronize(staged_[u])) != hipSuccess) return e;
        CHDesc* h = hDescs_[u];
        int tiles = 0, trs = 0, cols = 0;
        for (int i = 0; i < n; i++) {
            const CHRequest* r = active_[i];
            CHDesc& d = h[i];
            std::memset(&d, 0, sizeof(d));
            d.f[0] = r->buf.f[0];
            d.f[1] = r->buf.f[1];
            d.f[2] = r->buf.f[2];
            d.ring = 3;
            d.X = r->buf.X;
            d.Y = r->buf.Y;
            d.fac = r->buf.fac;
            d.a = r->co.a;
            d.b = r->co.b;
            d.d = r->co.d;
            d.e = r->co.e;
            for (int k = 0; k < 4; k++) d.omega[k] = r->co.omega[k];
            for (int k = 0; k < 25; k++) d.wlin[k] = r->co.wlin[k];
            for (int k = 0; k < 9; k++) d.wnl[k] = r->co.wnl[k];
            d.n = r->co.n;
            d.admitStep = r->admitStep;
            d.constFwd1 = r->co.constFwd + 1;
            d.constBwd1 = r->co.constBwd + 1;
            d.fastDiv = r->co.fastDiv;
            d.tileBase = tiles;
            d.colBase = cols;
            d.trBase = trs;
            tiles += chTiles(d.n);
            trs += chTrTiles(d.n);
            cols += chColBlocks(d.n);
        }
        if ((e = hipMemcpyAsync(dDescs_, h, sizeof(CHDesc) * n, hipMemcpyHostToDevice, st)) !=
                hipSuccess ||
            (e = hipEventRecord(staged_[u], st)) != hipSuccess)
            return e;
        stagedUsed_[u] = true;
        count_ = n;
        tiles_ = tiles;
        trTiles_ = trs;
        colBlocks_ = cols;
        dirty_ = false;
        version_++;
        return hipSuccess;
    }

    int device_;
    int step_ = 0;
    std::vector<CHRequest*> active_;
    bool dirty_ = true;
    CHDesc* dDescs_ = nullptr;
    CHDesc* hDescs_[2] = {nullptr, nullptr};  // pinned staging, used in turn
    hipEvent_t staged_[2] = {nullptr, nullptr};
    bool stagedUsed_[2] = {false, false};
    uint64_t uploads_ = 0;
    int cap_ = 0;
    int count_ = 0, tiles_ = 0, trTiles_ = 0, colBlocks_ = 0;
    // captured batches, used alternately
    struct BatchGraph {
        hipGraphExec_t exec = nullptr;
        int k = 0;
        int parity = 0;  // the batch's first step modulo 3
        uint64_t version = 0;
    };
    BatchGraph graphs_[2];
    uint64_t launches_ = 0;
    uint64_t version_ = 0;  // bumped by every descriptor upload
};

}  // namespace

bool chDescribe(const custenCHParams& p, CHDesc& d, double& cDiag, std::string& why)
{
    CHCoeffs co;
    if (!chConstants(p, co, why)) return false;
    std::memset(&d, 0, sizeof(d));
    d.a = co.a;
    d.b = co.b;
    d.d = co.d;
    d.e = co.e;
    for (int k = 0; k < 25; k++) d.wlin[k] = co.wlin[k];
    for (int k = 0; k < 9; k++) d.wnl[k] = co.wnl[k];
    d.n = co.n;
    cDiag = co.c;
    return true;
}

}  // namespace custen

using namespace custen;

// ===========================================================================
// Single simulation
// ===========================================================================
struct custenCH {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<CHEngine> engine;
    std::unique_ptr<BufferPool> pool;
    CHRequest req;
    // reference-shaped baseline (allocated on first use)
    PlainCH plain;
    double* wDev = nullptr;  // 25 linear + 9 nonlinear weights
    cuSten_t lin{}, nl{};
    bool linReady = false, nlReady = false;
    const double* nlIn = nullptr;
};

namespace {

void releasePlain(custenCH* ch)
{
    if (ch->linReady) cuStenDestroy2DXYp(&ch->lin);
    if (ch->nlReady) cuStenDestroy2DXYpFun(&ch->nl);
    ch->linReady = ch->nlReady = false;
    for (double* p : {ch->plain.cBar, ch->plain.nonLin, ch->plain.cHalf, ch->plain.multi, ch->wDev})
        if (p) (void)hipFree(p);
    ch->plain.cBar = ch->plain.nonLin = ch->plain.cHalf = ch->plain.multi = ch->wDev = nullptr;
}

hipError_t preparePlain(custenCH* ch)
{
    if (ch->plain.cBar) return hipSuccess;
    const CHCoeffs& co = ch->req.co;
    const int n = co.n, m = n - 2;
    const size_t nn = (size_t)n * n;
    PlainCH& p = ch->plain;
    p.n = n;
    p.a = co.a;
    p.b = co.b;
    p.d = co.d;
    p.e = co.e;
    for (int k = 0; k < 4; k++) p.omega[k] = co.omega[k];
    hipError_t e;
    if ((e = hipMalloc(&p.cBar, nn * 8)) != hipSuccess ||
        (e = hipMalloc(&p.nonLin, nn * 8)) != hipSuccess ||
        (e = hipMalloc(&p.cHalf, nn * 8)) != hipSuccess ||
        (e = hipMalloc(&p.multi, (size_t)7 * m * n * 8)) != hipSuccess ||
        (e = hipMalloc(&ch->wDev, 34 * 8)) != hipSuccess)
        return e;
    // replicate every factored diagonal and inverse across the n systems
    // (setMultiLHS + pentFactorBatch and the inv1Multi/inv2Multi loop of
    // cuPentCahnADI.cu:382-446 give these same values)
    std::vector<double> rep((size_t)7 * m * n);
    for (int a = 0; a < 7; a++)
        for (int i = 0; i < m; i++)
            std::fill_n(rep.begin() + ((size_t)a * m + i) * n, n, co.fac[(size_t)a * m + i]);
    if ((e = hipMemcpy(p.multi, rep.data(), rep.size() * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    double w[34];
    std::memcpy(w, co.wlin, sizeof(co.wlin));
    std::memcpy(w + 25, co.wnl, sizeof(co.wnl));
    return hipMemcpy(ch->wDev, w, sizeof(w), hipMemcpyHostToDevice);
}

}  // namespace

namespace {

custenCH* createSolver(const char* fn, int deviceNum, const custenCHParams* params,
                       const double* c0, const double* cOld0, int stepsDone)
{
    if (params == nullptr || c0 == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": null argument");
        return nullptr;
    }
    std::unique_ptr<custenCH> ch(new custenCH());
    std::string why;
    if (!chSetup(*params, ch->req.co, why)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": " + why);
        return nullptr;
    }
    hipError_t e = hipSetDevice(deviceNum);
    if (e != hipSuccess) {
        setHipError(e, fn);
        return nullptr;
    }
    ch->device = deviceNum;
    if ((e = hipStreamCreate(&ch->stream)) != hipSuccess) {
        setHipError(e, fn);
        return nullptr;
    }
    ch->engine.reset(new CHEngine(deviceNum));
    ch->pool.reset(new BufferPool(0));
    ch->req.params = *params;
    ch->req.done = stepsDone;
    if ((e = ch->pool->get(params->n, ch->req.buf)) != hipSuccess) {
        setHipError(e, fn);
        (void)hipStreamDestroy(ch->stream);
        return nullptr;
    }
    if ((e = ch->engine->admit(&ch->req, c0, cOld0, ch->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ch->stream)) != hipSuccess) {
        setHipError(e, fn);
        ch->pool->put(ch->req.buf);
        (void)hipStreamDestroy(ch->stream);
        return nullptr;
    }
    ch->req.state = 1;
    return ch.release();
}

// Checkpoint file: "CUSTENCH", u32 version, u32 n, f64 lx, D, gamma, dtOverDx,
// i64 steps done, then c and cOld as n*n little-endian doubles each.
constexpr char kCkptMagic[8] = {'C', 'U', 'S', 'T', 'E', 'N', 'C', 'H'};
constexpr uint32_t kCkptVersion = 1;
constexpr size_t kCkptHeader = 8 + 4 + 4 + 4 * 8 + 8;

bool writeFileAtomic(const std::string& path, const std::vector<std::pair<const void*, size_t>>& parts,
                     std::string& why)
{
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        why = "cannot open " + tmp + " for writing";
        return false;
    }
    bool ok = true;
    for (const auto& p : parts)
        ok = ok && std::fwrite(p.first, 1, p.second, f) == p.second;
    ok = (std::fclose(f) == 0) && ok;
    if (ok && std::rename(tmp.c_str(), path.c_str()) != 0) ok = false;
    if (!ok) {
        std::remove(tmp.c_str());
        why = "write to " + path + " failed";
    }
    return ok;
}

}  // namespace

extern "C" custenCH* custenCHCreate(int deviceNum, const custenCHParams* params, const double* c0)
{
    return createSolver("custenCHCreate", deviceNum, params, c0, nullptr, 0);
}

extern "C" int custenCHStepsDone(custenCH* ch)
{
    if (ch == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHStepsDone: null handle");
        return -1;
    }
    return ch->req.done;
}

extern "C" int custenCHCheckpoint(custenCH* ch, const char* path)
{
    if (ch == nullptr || path == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHCheckpoint: null argument");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    const int n = ch->req.co.n;
    const size_t nn = (size_t)n * n;
    std::vector<double> c(nn), cOld(nn);
    int rc = custenCHGetField(ch, c.data(), cOld.data());
    if (rc != CUSTEN_SUCCESS) return rc;
    const custenCHParams& pr = ch->req.params;
    const uint32_t ver = kCkptVersion, un = (uint32_t)n;
    const double prm[4] = {pr.lx, pr.D, pr.gamma, pr.dtOverDx};
    const int64_t done = ch->req.done;
    std::string why;
    if (!writeFileAtomic(path, {{kCkptMagic, 8}, {&ver, 4}, {&un, 4}, {prm, sizeof(prm)},
                                {&done, 8}, {c.data(), nn * 8}, {cOld.data(), nn * 8}},
                         why)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHCheckpoint: " + why);
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    return CUSTEN_SUCCESS;
}

extern "C" custenCH* custenCHRestore(int deviceNum, const char* path)
{
    const char* fn = "custenCHRestore";
    if (path == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": null path");
        return nullptr;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        setError(CUSTEN_ERROR_INVALID_VALUE, std::string(fn) + ": cannot open " + path);
        return nullptr;
    }
    char magic[8];
    uint32_t ver = 0, un = 0;
    double prm[4];
    int64_t done = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCkptMagic, 8) == 0 &&
              std::fread(&ver, 4, 1, f) == 1 && ver == kCkptVersion &&
              std::fread(&un, 4, 1, f) == 1 && un >= 8 && un <= (uint32_t)kCHMaxN &&
              std::fread(prm, sizeof(prm), 1, f) == 1 && std::fread(&done, 8, 1, f) == 1 &&
              done >= 0 && done <= 0x7fffffff;
    std::vector<double> c, cOld;
    if (ok) {
        const size_t nn = (size_t)un * un;
        c.resize(nn);
        cOld.resize(nn);
        ok = std::fread(c.data(), 8, nn, f) == nn && std::fread(cOld.data(), 8, nn, f) == nn &&
             std::fgetc(f) == EOF;
    }
    std::fclose(f);
    if (!ok) {
        setError(CUSTEN_ERROR_INVALID_VALUE,
                 std::string(fn) + ": " + path + " is not a complete checkpoint of this format");
        return nullptr;
    }
    custenCHParams pr{};
    pr.n = (int)un;
    pr.lx = prm[0];
    pr.D = prm[1];
    pr.gamma = prm[2];
    pr.dtOverDx = prm[3];
    return createSolver(fn, deviceNum, &pr, c.data(), cOld.data(), (int)done);
}

extern "C" int custenCHSnapshot(custenCH* ch, const char* path)
{
    if (ch == nullptr || path == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHSnapshot: null argument");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    const int n = ch->req.co.n;
    const size_t nn = (size_t)n * n;
    std::vector<double> c(nn);
    int rc = custenCHGetField(ch, c.data(), nullptr);
    if (rc != CUSTEN_SUCCESS) return rc;
    // NumPy .npy version 1.0: magic, version, header length, header dict
    // padded with spaces to a 64-byte boundary and ended by a newline
    std::string dict = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(n) +
                       ", " + std::to_string(n) + "), }";
    const size_t pre = 10;
    size_t total = pre + dict.size() + 1;
    dict.append((64 - total % 64) % 64, ' ');
    dict.push_back('\n');
    const uint16_t hlen = (uint16_t)dict.size();
    const unsigned char head[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::string why;
    if (!writeFileAtomic(path, {{head, 8}, {&hlen, 2}, {dict.data(), dict.size()},
                                {c.data(), nn * 8}},
                         why)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHSnapshot: " + why);
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    return CUSTEN_SUCCESS;
}

extern "C" int custenCHRun(custenCH* ch, int steps, int mode, hipStream_t stream)
{
    if (ch == nullptr || steps < 0 || (mode != 0 && mode != 1)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHRun: bad argument");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    (void)hipSetDevice(ch->device);
    hipStream_t st = stream ? stream : ch->stream;
    if (mode == 0) {
        hipError_t e = ch->engine->run(st, steps);
        if (e != hipSuccess) {
            setHipError(e, "custenCHRun");
            return lastError();
        }
        ch->req.done += steps;
        return CUSTEN_SUCCESS;
    }
    // mode 1: reference-shaped launch-and-synchronise loop
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = preparePlain(ch);
    if (e != hipSuccess) {
        setHipError(e, "custenCHRun");
        return lastError();
    }
    const int n = ch->req.co.n;
    auto stencil = [&](double* out, const double* in, bool nonlinear) -> hipError_t {
        if (nonlinear) {
            if (!ch->nlReady || ch->nlIn != in) {
                if (ch->nlReady) cuStenDestroy2DXYpFun(&ch->nl);
                double* f = custenGetBuiltinFunction("xy_cube_minus", ch->device);
                cuStenCreate2DXYpFun(&ch->nl, ch->device, 1, n, n, 8, 8, out,
                                     const_cast<double*>(in), ch->wDev + 25, 3, 1, 1, 3, 1, 1, f);
                if (lastError() != CUSTEN_SUCCESS) return hipErrorInvalidValue;
                custenSetStream(&ch->nl, st);
                ch->nlReady = true;
                ch->nlIn = in;
            }
            cuStenCompute2DXYpFun(&ch->nl, DEVICE);
        } else {
            if (!ch->linReady) {
                cuStenCreate2DXYp(&ch->lin, ch->device, 1, n, n, 32, 32, out,
                                  const_cast<double*>(in), ch->wDev, 5, 2, 2, 5, 2, 2);
                if (lastError() != CUSTEN_SUCCESS) return hipErrorInvalidValue;
                custenSetStream(&ch->lin, st);
                ch->linReady = true;
            }
            cuStenCompute2DXYp(&ch->lin, DEVICE);
        }
        return lastError() == CUSTEN_SUCCESS ? hipSuccess : hipErrorInvalidValue;
    };
    for (int s = 0; s < steps; s++) {
        double* cur = ch->engine->current(&ch->req);
        double* old = ch->engine->previous(&ch->req);
        e = chPlainStep(ch->plain, cur, old, st, stencil);
        if (e != hipSuccess) {
            if (lastError() == CUSTEN_SUCCESS) setHipError(e, "custenCHRun");
            return lastError();
        }
        ch->engine->noteExternalStep(&ch->req);
    }
    ch->req.done += steps;
    return CUSTEN_SUCCESS;
}

extern "C" int custenCHGetField(custenCH* ch, double* c, double* cOld)
{
    if (ch == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHGetField: null handle");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    (void)hipSetDevice(ch->device);
    const size_t bytes = (size_t)ch->req.co.n * ch->req.co.n * sizeof(double);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && c)
        e = hipMemcpy(c, ch->engine->current(&ch->req), bytes, hipMemcpyDefault);
    if (e == hipSuccess && cOld)
        e = hipMemcpy(cOld, ch->engine->previous(&ch->req), bytes, hipMemcpyDefault);
    if (e != hipSuccess) {
        setHipError(e, "custenCHGetField");
        return lastError();
    }
    return CUSTEN_SUCCESS;
}

extern "C" void custenCHDestroy(custenCH* ch)
{
    if (ch == nullptr) return;
    (void)hipSetDevice(ch->device);
    (void)hipStreamSynchronize(ch->stream);
    releasePlain(ch);
    ch->engine->remove(&ch->req);
    ch->pool->put(ch->req.buf);
    ch->engine.reset();
    ch->pool.reset();
    (void)hipStreamDestroy(ch->stream);
    delete ch;
}

extern "C" size_t custenCHFootprint(int n) { return n >= 8 ? footprintBytes(n) : 0; }

// ===========================================================================
// Server
// ===========================================================================
// Requests of one grid size share a lane: one batched engine on its own
// stream.  Lanes advance independently (a lane of small grids is not held to
// the pace of a lane of large ones), several steps per enqueue, with up to
// two batches in flight so that a lane keeps the GPU busy while the host
// attends to the others; the host only waits when no lane can take more
// work, and then for whichever batch finishes first.
struct CHLane {
    int n = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};  // end of each batch in flight
    int steps[2] = {0, 0};                     // steps of each batch in flight
    int head = 0;                              // oldest batch in flight
    int count = 0;                             // batches in flight
    std::unique_ptr<CHEngine> engine;
    int inFlight() const
    {
        int k = 0;
        for (int i = 0; i < count; i++) k += steps[(head + i) & 1];
        return k;
    }
};

// Pinned host buffers for results, reused by size.
class PinnedPool {
public:
    ~PinnedPool()
    {
        for (auto& kv : free_)
            for (double* p : kv.second) (void)hipHostFree(p);
    }
    hipError_t get(size_t count, double*& out)
    {
        auto it = free_.find(count);
        if (it != free_.end() && !it->second.empty()) {
            out = it->second.back();
            it->second.pop_back();
            return hipSuccess;
        }
        return hipHostMalloc(&out, count * sizeof(double), 0);
    }
    void put(double* p, size_t count) { free_[count].push_back(p); }

private:
    std::map<size_t, std::vector<double*>> free_;
};

// Results copied out of a lane's buffers by one retirement, and the event
// that marks the end of those copies.
struct CHDrain {
    hipEvent_t done = nullptr;
    std::vector<CHRequest*> reqs;
};

struct custenCHServer {
    int device = 0;
    size_t budget = 0;
    int maxSlots = 0;
    int resident = 0;
    size_t used = 0;      // device bytes of staged, resident and draining requests
    int unstaged = 0;     // queued requests whose input waits in host memory
    int draining = 0;     // requests whose result is being copied out
    size_t pinnedBytes = 0;  // results held in pinned memory
    size_t pinnedCap = 0;    // ... up to this many bytes
    hipStream_t upload = nullptr;
    std::map<int, std::unique_ptr<CHLane>> lanes;
    std::unique_ptr<BufferPool> pool;
    PinnedPool pinned;
    std::deque<CHRequest*> queue;
    std::vector<CHDrain> drains;
    std::vector<std::unique_ptr<CHRequest>> all;
};

namespace {

// Steps a lane enqueues at once (fewer if a request finishes earlier).
constexpr int kLaneBatch = 8;
// Finished results kept in pinned memory up to this many bytes (the
// environment variable CUSTEN_CH_PINNED_RESULT_BYTES, read when a server is
// created, overrides it); later ones move to pageable memory once their copy
// has ended.
constexpr size_t kPinnedResultBytes = size_t(2) << 30;

void releaseLane(CHLane& l)
{
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    l.engine.reset();
    for (hipEvent_t& ev : l.done) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    if (l.stream) (void)hipStreamDestroy(l.stream);
    l.stream = nullptr;
}

hipError_t laneFor(custenCHServer* s, int n, CHLane*& out)
{
    auto it = s->lanes.find(n);
    if (it != s->lanes.end()) {
        out = it->second.get();
        return hipSuccess;
    }
    std::unique_ptr<CHLane> l(new CHLane());
    l->n = n;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking)) != hipSuccess) return e;
    for (hipEvent_t& ev : l->done) {
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            releaseLane(*l);
            return e;
        }
    }
    l->engine.reset(new CHEngine(s->device));
    out = l.get();
    s->lanes[n] = std::move(l);
    return hipSuccess;
}

}  // namespace

extern "C" custenCHServer* custenCHServerCreate(int deviceNum, size_t hbmBudgetBytes, int maxSlots)
{
    if (hbmBudgetBytes == 0 || maxSlots < 1) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerCreate: budget and slots must be > 0");
        return nullptr;
    }
    hipError_t e = hipSetDevice(deviceNum);
    if (e != hipSuccess) {
        setHipError(e, "custenCHServerCreate");
        return nullptr;
    }
    std::unique_ptr<custenCHServer> s(new custenCHServer());
    s->device = deviceNum;
    s->budget = hbmBudgetBytes;
    s->maxSlots = maxSlots;
    s->pool.reset(new BufferPool(hbmBudgetBytes / 4));
    s->pinnedCap = kPinnedResultBytes;
    if (const char* v = std::getenv("CUSTEN_CH_PINNED_RESULT_BYTES"))
        s->pinnedCap = (size_t)std::strtoull(v, nullptr, 10);
    if ((e = hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking)) != hipSuccess) {
        setHipError(e, "custenCHServerCreate");
        return nullptr;
    }
    return s.release();
}

extern "C" int custenCHServerSubmitResume(custenCHServer* s, const custenCHParams* params,
                                          int steps, const double* c0, const double* cOld0)
{
    if (s == nullptr || params == nullptr || c0 == nullptr || steps < 1) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerSubmit: bad argument");
        return -1;
    }
    std::unique_ptr<CHRequest> r(new CHRequest());
    std::string why;
    if (!chSetup(*params, r->co, why)) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerSubmit: " + why);
        return -1;
    }
    if (footprintBytes(params->n) > s->budget) {
        setError(CUSTEN_ERROR_ADMISSION,
                 "custenCHServerSubmit: request footprint exceeds the HBM budget");
        return -1;
    }
    r->id = (int)s->all.size();
    r->params = *params;
    r->steps = steps;
    r->tSubmit = Clock::now();
    // Stage the input in HBM now when the budget has room and every request
    // ahead of this one is staged too (so staged requests always head the
    // queue and cannot hold memory that an earlier request needs); else keep
    // a host copy until admission.
    const size_t fp = footprintBytes(params->n);
    (void)hipSetDevice(s->device);
    bool stage = s->unstaged == 0 && s->used + fp <= s->budget;
    if (stage && s->pool->get(params->n, r->buf) != hipSuccess) {
        (void)hipGetLastError();  // a failed allocation only means "not now"
        stage = false;
    }
    if (stage) {
        hipError_t e = loadRequest(r.get(), c0, cOld0, s->upload);
        if (e == hipSuccess) e = hipStreamSynchronize(s->upload);
        if (e != hipSuccess) {
            s->pool->put(r->buf);
            setHipError(e, "custenCHServerSubmit");
            return -1;
        }
        r->staged = true;
        s->used += fp;
    } else {
        const size_t nn = (size_t)params->n * params->n;
        r->c0.assign(c0, c0 + nn);
        if (cOld0) r->cOld0.assign(cOld0, cOld0 + nn);
        s->unstaged++;
    }
    s->queue.push_back(r.get());
    s->all.push_back(std::move(r));
    return s->all.back()->id;
}

extern "C" int custenCHServerSubmit(custenCHServer* s, const custenCHParams* params, int steps,
                                    const double* c0)
{
    return custenCHServerSubmitResume(s, params, steps, c0, nullptr);
}

namespace {

// A request whose result has reached host memory: its device buffers go back
// to the pool.
void finishRequest(custenCHServer* s, CHRequest* r)
{
    s->pool->put(r->buf);
    r->buf = CHBuffers{};
    s->used -= footprintBytes(r->co.n);
    std::vector<double>().swap(r->c0);
    std::vector<double>().swap(r->cOld0);
    r->tFinish = Clock::now();
    r->state = 2;
}

// Account a lane's oldest batch, which has finished: every resident request
// advanced by the batch's steps; finished ones leave the batch and their
// results are copied to pinned host memory behind it on the lane's stream,
// ending in a drain event (a batch never runs past a request's last step,
// and a lane with a request about to finish has one batch in flight, so the
// copies follow the request's last step and precede any later batch).
hipError_t retireBatch(custenCHServer* s, CHLane& l)
{
    hipError_t e;
    const int k = l.steps[l.head];
    l.steps[l.head] = 0;
    l.head ^= 1;
    l.count--;
    std::vector<CHRequest*> fin;
    for (CHRequest* r : l.engine->active()) {
        r->done += k;
        if (r->done >= r->steps) fin.push_back(r);
    }
    if (fin.empty()) return hipSuccess;
    CHDrain d;
    for (CHRequest* r : fin) {
        const size_t nn = (size_t)r->co.n * r->co.n;
        if ((e = s->pinned.get(nn, r->pinnedResult)) != hipSuccess) {
            // no pinned memory: copy to pageable memory and wait
            (void)hipGetLastError();
            r->pinnedResult = nullptr;
            r->result.resize(nn);
            if ((e = hipMemcpyAsync(r->result.data(), l.engine->current(r), nn * sizeof(double),
                                    hipMemcpyDeviceToHost, l.stream)) != hipSuccess ||
                (e = hipStreamSynchronize(l.stream)) != hipSuccess)
                return e;
            l.engine->remove(r);
            s->resident--;
            finishRequest(s, r);
            continue;
        }
        s->pinnedBytes += nn * sizeof(double);
        if ((e = hipMemcpyAsync(r->pinnedResult, l.engine->current(r), nn * sizeof(double),
                                hipMemcpyDeviceToHost, l.stream)) != hipSuccess)
            return e;
        l.engine->remove(r);
        s->resident--;
        s->draining++;
        d.reqs.push_back(r);
    }
    if (d.reqs.empty()) return hipSuccess;
    if ((e = hipEventCreateWithFlags(&d.done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventRecord(d.done, l.stream)) != hipSuccess)
        return e;
    s->drains.push_back(std::move(d));
    return hipSuccess;
}

// Finish the requests of every drain whose copies have ended.  Returns the
// number of requests finished, or -1 on error.
int finishDrains(custenCHServer* s)
{
    int finished = 0;
    for (size_t i = 0; i < s->drains.size();) {
        CHDrain& d = s->drains[i];
        const hipError_t q = hipEventQuery(d.done);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        if (q != hipSuccess) {
            setHipError(q, "custenCHServerStep (result copy)");
            return -1;
        }
        for (CHRequest* r : d.reqs) {
            finishRequest(s, r);
            s->draining--;
            finished++;
            if (s->pinnedBytes > s->pinnedCap) {
                const size_t nn = (size_t)r->co.n * r->co.n;
                r->result.assign(r->pinnedResult, r->pinnedResult + nn);
                s->pinned.put(r->pinnedResult, nn);
                r->pinnedResult = nullptr;
                s->pinnedBytes -= nn * sizeof(double);
            }
        }
        (void)hipEventDestroy(d.done);
        s->drains.erase(s->drains.begin() + i);
    }
    return finished;
}

// Retire every finished batch of every lane, and finish every ended drain,
// without waiting.  Returns the number retired, or -1 on error.
int retireFinished(custenCHServer* s)
{
    int retired = finishDrains(s);
    if (retired < 0) return -1;
    for (auto& kv : s->lanes) {
        CHLane& l = *kv.second;
        while (l.count > 0) {
            const hipError_t q = hipEventQuery(l.done[l.head]);
            if (q == hipErrorNotReady) break;
            hipError_t e = q;
            if (e == hipSuccess) e = retireBatch(s, l);
            if (e != hipSuccess) {
                setHipError(e, "custenCHServerStep (retire)");
                return -1;
            }
            retired++;
        }
    }
    return retired;
}

// One scheduling round: retire the lanes' finished batches, admit queued
// requests (FIFO, while a slot and the HBM budget allow and the request's
// lane is idle), and enqueue the next batch of steps on every lane with
// resident requests and fewer than two batches in flight.  If nothing could
// be done without waiting, first wait for whichever batch in flight finishes
// first.
int serverIteration(custenCHServer* s)
{
    (void)hipSetDevice(s->device);
    hipError_t e;
    for (int attempt = 0;; attempt++) {
        const int retired = retireFinished(s);
        if (retired < 0) return -1;
        // admission: FIFO while a slot and the HBM budget allow
        int admitted = 0;
        while (!s->queue.empty() && s->resident < s->maxSlots) {
            CHRequest* r = s->queue.front();
            const size_t fp = footprintBytes(r->co.n);
            if (!r->staged && s->used + fp > s->budget) break;
            CHLane* l = nullptr;
            if ((e = laneFor(s, r->co.n, l)) != hipSuccess) {
                setHipError(e, "custenCHServerStep (lane)");
                return -1;
            }
            if (l->count > 0) break;  // its lane is mid-batch: admit at a later round
            if (r->staged) {
                l->engine->adopt(r);
            } else {
                if ((e = s->pool->get(r->co.n, r->buf)) != hipSuccess) {
                    if (s->resident == 0 && s->draining == 0) {
                        setHipError(e, "custenCHServerStep (allocation)");
                        return -1;
                    }
                    (void)hipGetLastError();
                    break;  // wait for memory to free up
                }
                if ((e = l->engine->admit(r, r->c0.data(),
                                          r->cOld0.empty() ? nullptr : r->cOld0.data(),
                                          l->stream)) != hipSuccess) {
                    setHipError(e, "custenCHServerStep (admit)");
                    return -1;
                }
                s->used += fp;
                s->unstaged--;
            }
            s->resident++;
            r->state = 1;
            r->tAdmit = Clock::now();
            s->queue.pop_front();
            admitted++;
        }
        // enqueue: a lane takes a batch while it has fewer than two in flight
        // and every resident request has steps left beyond those in flight;
        // a lane whose queue head waits for it to drain takes no second batch.
        // Largest grids first: their lane sets the pace of a mixed load.
        int enqueued = 0;
        for (auto it = s->lanes.rbegin(); it != s->lanes.rend(); ++it) {
            CHLane& l = *it->second;
            if (l.count >= 2 || l.engine->active().empty()) continue;
            if (l.count > 0 && !s->queue.empty() && s->queue.front()->co.n == l.n) continue;
            const int flight = l.inFlight();
            int k = kLaneBatch;
            for (CHRequest* r : l.engine->active()) k = std::min(k, r->steps - r->done - flight);
            if (k < 1) continue;
            if ((e = l.engine->stepBatch(l.stream, k)) != hipSuccess) {
                setHipError(e, "custenCHServerStep");
                return -1;
            }
            const int slot = (l.head + l.count) & 1;
            if ((e = hipEventRecord(l.done[slot], l.stream)) != hipSuccess) {
                setHipError(e, "custenCHServerStep");
                return -1;
            }
            l.steps[slot] = k;
            l.count++;
            enqueued++;
        }
        const int left = (int)s->queue.size() + s->resident + s->draining;
        if (retired + admitted + enqueued > 0 || left == 0 || attempt > 0) return left;
        // nothing to do without waiting: wait for the first batch or result
        // copy to finish
        bool anyInFlight = !s->drains.empty();
        for (auto& kv : s->lanes) anyInFlight = anyInFlight || kv.second->count > 0;
        if (!anyInFlight) return left;
        for (;;) {
            bool ready = false;
            for (auto& kv : s->lanes) {
                CHLane& l = *kv.second;
                if (l.count > 0 && hipEventQuery(l.done[l.head]) != hipErrorNotReady) ready = true;
            }
            for (const CHDrain& d : s->drains)
                if (hipEventQuery(d.done) != hipErrorNotReady) ready = true;
            if (ready) break;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
}

}  // namespace

extern "C" int custenCHServerStep(custenCHServer* s)
{
    if (s == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerStep: null server");
        return -1;
    }
    return serverIteration(s);
}

extern "C" int custenCHServerRunUntilIdle(custenCHServer* s, int maxIterations)
{
    if (s == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerRunUntilIdle: null server");
        return -1;
    }
    int it = 0;
    while (maxIterations <= 0 || it < maxIterations) {
        const int left = serverIteration(s);
        if (left < 0) return -1;
        it++;
        if (left == 0) break;
    }
    return it;
}

extern "C" int custenCHServerQuery(custenCHServer* s, int id, int* state, double* latencySubmit,
                                   double* latencyAdmit)
{
    if (s == nullptr || id < 0 || id >= (int)s->all.size()) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerQuery: unknown request");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    const CHRequest* r = s->all[id].get();
    if (state) *state = r->state;
    if (r->state == 2) {
        if (latencySubmit)
            *latencySubmit = std::chrono::duration<double>(r->tFinish - r->tSubmit).count();
        if (latencyAdmit)
            *latencyAdmit = std::chrono::duration<double>(r->tFinish - r->tAdmit).count();
    } else {
        if (latencySubmit) *latencySubmit = -1.0;
        if (latencyAdmit) *latencyAdmit = -1.0;
    }
    return CUSTEN_SUCCESS;
}

namespace {

// Host copy of a result into the caller's array.  A large copy into memory
// the caller has not touched yet is bound by page faults, which the kernel
// serves in parallel: it is split over up to 8 threads.
void copyOut(double* dst, const double* src, size_t count)
{
    constexpr size_t kSplitBytes = size_t(8) << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned parts = std::max(1u, std::min(8u, hw));
    if (count * sizeof(double) < kSplitBytes || parts == 1) {
        std::memcpy(dst, src, count * sizeof(double));
        return;
    }
    const size_t chunk = (count + parts - 1) / parts;
    std::vector<std::thread> workers;
    for (unsigned t = 1; t < parts && t * chunk < count; t++) {
        const size_t b = t * chunk, e = std::min(count, b + chunk);
        workers.emplace_back([=]() { std::memcpy(dst + b, src + b, (e - b) * sizeof(double)); });
    }
    std::memcpy(dst, src, std::min(chunk, count) * sizeof(double));
    for (std::thread& w : workers) w.join();
}

}  // namespace

extern "C" int custenCHServerResult(custenCHServer* s, int id, double* c)
{
    if (s == nullptr || id < 0 || id >= (int)s->all.size() || c == nullptr) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerResult: bad argument");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    const CHRequest* r = s->all[id].get();
    if (r->state != 2) {
        setError(CUSTEN_ERROR_INVALID_VALUE, "custenCHServerResult: request not finished");
        return CUSTEN_ERROR_INVALID_VALUE;
    }
    const size_t nn = (size_t)r->co.n * r->co.n;
    copyOut(c, r->pinnedResult ? r->pinnedResult : r->result.data(), nn);
    return CUSTEN_SUCCESS;
}

extern "C" void custenCHServerDestroy(custenCHServer* s)
{
    if (s == nullptr) return;
    (void)hipSetDevice(s->device);
    for (auto& kv : s->lanes) {
        CHLane& l = *kv.second;
        (void)hipStreamSynchronize(l.stream);
        for (CHRequest* r : std::vector<CHRequest*>(l.engine->active())) {
            l.engine->remove(r);
            s->pool->put(r->buf);
            r->buf = CHBuffers{};
        }
        releaseLane(l);
    }
    s->lanes.clear();
    for (CHDrain& d : s->drains) {
        (void)hipEventSynchronize(d.done);
        (void)hipEventDestroy(d.done);
        for (CHRequest* r : d.reqs) {
            s->pool->put(r->buf);
            r->buf = CHBuffers{};
        }
    }
    s->drains.clear();
    for (CHRequest* r : s->queue)
        if (r->staged) {
            s->pool->put(r->buf);
            r->buf = CHBuffers{};
        }
    for (auto& r : s->all)
        if (r->pinnedResult) {
            s->pinned.put(r->pinnedResult, (size_t)r->co.n * r->co.n);
            r->pinnedResult = nullptr;
        }
    if (s->upload) (void)hipStreamDestroy(s->upload);
    s->pool.reset();
    delete s;
}
