// tuner.cpp -- autotuner for the stencil tile configuration (rows per wave x
// row chunks per workgroup).  All configurations compute bit-identical
// results; the tuner times each candidate on the caller's stream with HIP
// events and keeps the fastest per (dtype, stencil shape, grid shape) in a
// process-wide table, optionally persisted in the file named by
// CUSTEN_TUNE_CACHE.  CUSTEN_AUTOTUNE=0 disables tuning (defaults are used).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "tuner.h"

namespace custen {
namespace {

std::mutex g_mu;
std::map<TuneKey, TuneVal> g_table;
bool g_loaded = false;

const char* cachePath()
{
    const char* p = std::getenv("CUSTEN_TUNE_CACHE");
    return (p && *p) ? p : nullptr;
}

void loadLocked()
{
    if (g_loaded) return;
    g_loaded = true;
    const char* path = cachePath();
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    TuneKey k;
    TuneVal v;
    while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d %d %d", &k.kind, &k.dtype, &k.H, &k.V,
                       &k.xform, &k.nx, &k.ny, &k.count, &v.rpw, &v.chunks, &v.cpl) == 11)
        g_table[k] = v;
    std::fclose(f);
}

}  // namespace

bool TuneKey::operator<(const TuneKey& o) const
{
    return std::tie(kind, dtype, H, V, xform, nx, ny, count) <
           std::tie(o.kind, o.dtype, o.H, o.V, o.xform, o.nx, o.ny, o.count);
}

bool tunerEnabled()
{
    const char* e = std::getenv("CUSTEN_AUTOTUNE");
    return !(e && std::strcmp(e, "0") == 0);
}

bool tunerLookup(const TuneKey& k, TuneVal& v)
{
    std::lock_guard<std::mutex> lock(g_mu);
    loadLocked();
    auto it = g_table.find(k);
    if (it == g_table.end()) return false;
    v = it->second;
    return true;
}

void tunerStore(const TuneKey& k, const TuneVal& v)
{
    std::lock_guard<std::mutex> lock(g_mu);
    loadLocked();
    g_table[k] = v;
    const char* path = cachePath();
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d\n", k.kind, k.dtype, k.H, k.V, k.xform, k.nx,
                 k.ny, k.count, v.rpw, v.chunks, v.cpl);
    std::fclose(f);
}

hipError_t tunerMeasure(hipStream_t st, const std::function<hipError_t(const TuneVal&)>& run,
                        const std::vector<TuneVal>& cands, int reps, TuneVal& best)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return e;
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    float bestMs = 1e30f;
    best = cands.front();
    for (const TuneVal& c : cands) {
        if ((e = run(c)) != hipSuccess) break;  // warm-up
        if ((e = hipEventRecord(e0, st)) != hipSuccess) break;
        for (int r = 0; r < reps && e == hipSuccess; r++) e = run(c);
        if (e != hipSuccess) break;
        if ((e = hipEventRecord(e1, st)) != hipSuccess) break;
        if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        if (ms < bestMs) {
            bestMs = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e;
}

bool streamCapturing(hipStream_t st)
{
    hipStreamCaptureStatus s = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &s) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // be conservative: do not time inside an unknown state
    }
    return s != hipStreamCaptureStatusNone;
}

}  // namespace custen
