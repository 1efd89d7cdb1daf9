// tuner.h -- tile-configuration autotuner (see tuner.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <functional>
#include <vector>

namespace custen {

struct TuneKey {
    int kind = 0;   // 0 single grid, 1 batched grids
    int dtype = 0;
    int H = 0, V = 0;
    int xform = 0;
    int nx = 0, ny = 0;
    int count = 0;  // batched: number of grids (nx, ny: total points / 2^20, max grid side)
    bool operator<(const TuneKey& o) const;
};

struct TuneVal {
    int rpw = 8;
    int chunks = 4;
    int cpl = 2;  // columns per lane
    int band = 0;  // 1: band walker (chunks = chunks per band)
};

bool tunerEnabled();
bool tunerLookup(const TuneKey& k, TuneVal& v);
void tunerStore(const TuneKey& k, const TuneVal& v);
// Times every candidate on `st` (at least `reps` launches each, more for
// short kernels, after a warm-up), re-times the three fastest, and leaves the
// fastest in `best`.
hipError_t tunerMeasure(hipStream_t st, const std::function<hipError_t(const TuneVal&)>& run,
                        const std::vector<TuneVal>& cands, int reps, TuneVal& best);
bool streamCapturing(hipStream_t st);

}  // namespace custen
