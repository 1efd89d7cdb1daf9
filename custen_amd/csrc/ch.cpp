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
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
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
    return true;
}

// Device buffers of one request, recycled by grid size.
struct CHBuffers {
    int n = 0;
    double* f[2] = {nullptr, nullptr};
    double* X = nullptr;
    double* Y = nullptr;
    double* fac = nullptr;
};

size_t footprintBytes(int n)
{
    const size_t nn = (size_t)n * n;
    return 4 * nn * sizeof(double) + (size_t)7 * (n - 2) * sizeof(double);
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
        for (double* p : {b.f[0], b.f[1], b.X, b.Y, b.fac})
            if (p) (void)hipFree(p);
        b.f[0] = b.f[1] = b.X = b.Y = b.fac = nullptr;
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
    std::vector<double> result;
    int state = 0;  // 0 queued, 1 resident, 2 finished
    Clock::time_point tSubmit, tAdmit, tFinish;
};

// Advances every resident request by one step with one launch per phase.
class CHEngine {
public:
    explicit CHEngine(int device) : device_(device) {}
    ~CHEngine()
    {
        if (dDescs_) (void)hipFree(dDescs_);
        if (hDescs_) (void)hipHostFree(hDescs_);
    }

    int step() const { return step_; }

    // cOld0 == nullptr: cOld starts equal to c (cuPentCahnADI.cu:295-305);
    // otherwise the pair (c0, cOld0) resumes a simulation exactly.
    hipError_t admit(CHRequest* r, const double* c0, const double* cOld0, hipStream_t st)
    {
        const size_t bytes = (size_t)r->co.n * r->co.n * sizeof(double);
        hipError_t e;
        if ((e = hipMemcpyAsync(r->buf.f[0], c0, bytes, hipMemcpyDefault, st)) != hipSuccess) return e;
        if (cOld0) e = hipMemcpyAsync(r->buf.f[1], cOld0, bytes, hipMemcpyDefault, st);
        else e = hipMemcpyAsync(r->buf.f[1], r->buf.f[0], bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        if ((e = hipMemcpyAsync(r->buf.fac, r->co.fac.data(), r->co.fac.size() * sizeof(double),
                                hipMemcpyHostToDevice, st)) != hipSuccess)
            return e;
        r->admitStep = step_;
        active_.push_back(r);
        dirty_ = true;
        return hipSuccess;
    }

    void remove(CHRequest* r)
    {
        active_.erase(std::remove(active_.begin(), active_.end(), r), active_.end());
        dirty_ = true;
    }

    const std::vector<CHRequest*>& active() const { return active_; }

    // Current field of a resident request.
    double* current(const CHRequest* r) const { return r->buf.f[(step_ - r->admitStep) & 1]; }
    double* previous(const CHRequest* r) const
    {
        return r->buf.f[((step_ - r->admitStep) & 1) ^ 1];
    }

    // A step done outside the engine (the plain path) that kept the field in
    // place: advance the step counter without flipping the request's roles.
    void noteExternalStep(CHRequest* r)
    {
        step_++;
        r->admitStep++;
        dirty_ = true;
    }

    // One step of all resident requests.  plain = true synchronises the stream
    // after every launch.
    hipError_t stepAll(hipStream_t st, bool plain)
    {
        if (active_.empty()) return hipSuccess;
        hipError_t e;
        if (dirty_ && (e = upload(st)) != hipSuccess) return e;
        auto sync = [&]() { return plain ? hipStreamSynchronize(st) : hipSuccess; };
        if ((e = chLaunchRhs(dDescs_, count_, tiles_, step_, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        if ((e = chLaunchTranspose(dDescs_, count_, trTiles_, 0, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        if ((e = chLaunchColSolve(dDescs_, count_, colBlocks_, 0, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        if ((e = chLaunchTranspose(dDescs_, count_, trTiles_, 1, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        if ((e = chLaunchColSolve(dDescs_, count_, colBlocks_, 1, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        if ((e = chLaunchFinal(dDescs_, count_, tiles_, step_, st)) != hipSuccess) return e;
        if ((e = sync()) != hipSuccess) return e;
        step_++;
        return hipSuccess;
    }

private:
    hipError_t upload(hipStream_t st)
    {
        const int n = (int)active_.size();
        if (n > cap_) {
            // the previous descriptor upload may still be read by queued launches
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            if (dDescs_) (void)hipFree(dDescs_);
            if (hDescs_) (void)hipHostFree(hDescs_);
            dDescs_ = nullptr;
            hDescs_ = nullptr;
            cap_ = std::max(n, 2 * cap_);
            if ((e = hipMalloc(&dDescs_, sizeof(CHDesc) * cap_)) != hipSuccess) return e;
            if ((e = hipHostMalloc(&hDescs_, sizeof(CHDesc) * cap_, 0)) != hipSuccess) return e;
        } else {
            // pinned staging is reused: wait until the previous copy consumed it
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
        }
        int tiles = 0, trs = 0, cols = 0;
        for (int i = 0; i < n; i++) {
            const CHRequest* r = active_[i];
            CHDesc& d = hDescs_[i];
            std::memset(&d, 0, sizeof(d));
            d.f[0] = r->buf.f[0];
            d.f[1] = r->buf.f[1];
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
            d.tileBase = tiles;
            d.colBase = cols;
            d.trBase = trs;
            tiles += chTiles(d.n);
            trs += chTrTiles(d.n);
            cols += chColBlocks(d.n);
        }
        hipError_t e = hipMemcpyAsync(dDescs_, hDescs_, sizeof(CHDesc) * n, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        count_ = n;
        tiles_ = tiles;
        trTiles_ = trs;
        colBlocks_ = cols;
        dirty_ = false;
        return hipSuccess;
    }

    int device_;
    int step_ = 0;
    std::vector<CHRequest*> active_;
    bool dirty_ = true;
    CHDesc* dDescs_ = nullptr;
    CHDesc* hDescs_ = nullptr;
    int cap_ = 0;
    int count_ = 0, tiles_ = 0, trTiles_ = 0, colBlocks_ = 0;
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
        for (int s = 0; s < steps; s++) {
            hipError_t e = ch->engine->stepAll(st, false);
            if (e != hipSuccess) {
                setHipError(e, "custenCHRun");
                return lastError();
            }
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
// the pace of a lane of large ones), several steps per enqueue; the host only
// waits when no lane can take more work.
struct CHLane {
    int n = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::unique_ptr<CHEngine> engine;
    int inFlight = 0;            // steps enqueued and not yet accounted
    Clock::time_point enqueued;  // when the in-flight batch was enqueued
};

struct custenCHServer {
    int device = 0;
    size_t budget = 0;
    int maxSlots = 0;
    int resident = 0;
    size_t used = 0;
    std::map<int, std::unique_ptr<CHLane>> lanes;
    std::unique_ptr<BufferPool> pool;
    std::deque<CHRequest*> queue;
    std::vector<std::unique_ptr<CHRequest>> all;
};

namespace {

// Steps a lane enqueues at once (fewer if a request finishes earlier).
constexpr int kLaneBatch = 8;

void releaseLane(CHLane& l)
{
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    l.engine.reset();
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
    l.done = nullptr;
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
    if ((e = hipEventCreateWithFlags(&l->done, hipEventDisableTiming)) != hipSuccess) {
        releaseLane(*l);
        return e;
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
    const size_t nn = (size_t)params->n * params->n;
    r->c0.assign(c0, c0 + nn);
    if (cOld0) r->cOld0.assign(cOld0, cOld0 + nn);
    r->tSubmit = Clock::now();
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

// Account a lane's finished batch: every resident request advanced by the
// batch's steps; finished ones are copied out and retired.
hipError_t retireLane(custenCHServer* s, CHLane& l)
{
    hipError_t e;
    std::vector<CHRequest*> fin;
    for (CHRequest* r : l.engine->active()) {
        r->done += l.inFlight;
        if (r->done >= r->steps) fin.push_back(r);
    }
    l.inFlight = 0;
    for (CHRequest* r : fin) {
        const size_t nn = (size_t)r->co.n * r->co.n;
        r->result.resize(nn);
        if ((e = hipMemcpyAsync(r->result.data(), l.engine->current(r), nn * sizeof(double),
                                hipMemcpyDeviceToHost, l.stream)) != hipSuccess ||
            (e = hipStreamSynchronize(l.stream)) != hipSuccess)
            return e;
        r->tFinish = Clock::now();
        r->state = 2;
        l.engine->remove(r);
        s->pool->put(r->buf);
        r->buf = CHBuffers{};
        s->used -= footprintBytes(r->co.n);
        s->resident--;
        std::vector<double>().swap(r->c0);
        std::vector<double>().swap(r->cOld0);
    }
    return hipSuccess;
}

// One scheduling round: retire the work of every lane whose batch finished,
// admit queued requests (FIFO, while a slot and the HBM budget allow and the
// request's lane is idle), enqueue the next batch of steps on every idle lane
// with resident requests.  If no lane was idle, first wait for the batch that
// was enqueued earliest.
int serverIteration(custenCHServer* s)
{
    (void)hipSetDevice(s->device);
    hipError_t e;
    auto idle = [](CHLane& l) {
        if (l.inFlight == 0) return true;
        const hipError_t q = hipEventQuery(l.done);
        if (q == hipErrorNotReady) return false;
        return true;  // finished, or an error that the next call reports
    };
    // can this round do anything without waiting?
    bool canWork = false;
    bool anyInFlight = false;
    for (auto& kv : s->lanes) {
        CHLane& l = *kv.second;
        anyInFlight = anyInFlight || l.inFlight > 0;
        if (l.inFlight > 0 ? idle(l) : !l.engine->active().empty()) canWork = true;
    }
    if (!canWork && !s->queue.empty() && s->resident < s->maxSlots) {
        const CHRequest* r = s->queue.front();
        auto it = s->lanes.find(r->co.n);
        if (s->used + footprintBytes(r->co.n) <= s->budget &&
            (it == s->lanes.end() || it->second->inFlight == 0))
            canWork = true;
    }
    if (!canWork && anyInFlight) {
        CHLane* first = nullptr;
        for (auto& kv : s->lanes)
            if (kv.second->inFlight > 0 && (!first || kv.second->enqueued < first->enqueued))
                first = kv.second.get();
        if (first && (e = hipEventSynchronize(first->done)) != hipSuccess) {
            setHipError(e, "custenCHServerStep");
            return -1;
        }
    }
    for (auto& kv : s->lanes) {
        CHLane& l = *kv.second;
        if (l.inFlight == 0 || !idle(l)) continue;
        if ((e = hipEventSynchronize(l.done)) != hipSuccess ||
            (e = retireLane(s, l)) != hipSuccess) {
            setHipError(e, "custenCHServerStep (retire)");
            return -1;
        }
    }
    // admission: FIFO while a slot and the HBM budget allow
    while (!s->queue.empty() && s->resident < s->maxSlots) {
        CHRequest* r = s->queue.front();
        const size_t fp = footprintBytes(r->co.n);
        if (s->used + fp > s->budget) break;
        CHLane* l = nullptr;
        if ((e = laneFor(s, r->co.n, l)) != hipSuccess) {
            setHipError(e, "custenCHServerStep (lane)");
            return -1;
        }
        if (l->inFlight > 0) break;  // its lane is mid-batch: admit at the next round
        if ((e = s->pool->get(r->co.n, r->buf)) != hipSuccess) {
            if (s->resident == 0) {
                setHipError(e, "custenCHServerStep (allocation)");
                return -1;
            }
            break;  // wait for memory to free up
        }
        if ((e = l->engine->admit(r, r->c0.data(), r->cOld0.empty() ? nullptr : r->cOld0.data(),
                                  l->stream)) != hipSuccess) {
            setHipError(e, "custenCHServerStep (admit)");
            return -1;
        }
        s->used += fp;
        s->resident++;
        r->state = 1;
        r->tAdmit = Clock::now();
        s->queue.pop_front();
    }
    // enqueue the next batch on every idle lane
    for (auto& kv : s->lanes) {
        CHLane& l = *kv.second;
        if (l.inFlight > 0 || l.engine->active().empty()) continue;
        int k = kLaneBatch;
        for (CHRequest* r : l.engine->active()) k = std::min(k, r->steps - r->done);
        k = std::max(k, 1);
        for (int i = 0; i < k; i++) {
            if ((e = l.engine->stepAll(l.stream, false)) != hipSuccess) {
                setHipError(e, "custenCHServerStep");
                return -1;
            }
        }
        if ((e = hipEventRecord(l.done, l.stream)) != hipSuccess) {
            setHipError(e, "custenCHServerStep");
            return -1;
        }
        l.inFlight = k;
        l.enqueued = Clock::now();
    }
    return (int)s->queue.size() + s->resident;
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
    std::memcpy(c, r->result.data(), r->result.size() * sizeof(double));
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
        }
        releaseLane(l);
    }
    s->lanes.clear();
    s->pool.reset();
    delete s;
}
