// autotune.hpp -- per-group launch-schedule tuning of batch plans
// (autotune.cpp).
#pragma once
#include <string>

#include "engine.hpp"

namespace mfx {

// one batch group: `count` entries of `kind` with dims n[0..d-1]; the
// scratch tables are the group's own (their first entries are compact)
struct TuneRequest {
    Prec prec = PREC_F32;
    int kind = 0;
    int d = 1;
    int n[3] = {1, 1, 1};
    unsigned count = 0;
    const long long* d_foff = nullptr;
    void* tbuf = nullptr;
    const long long* d_toff = nullptr;
    // the group is the plan's only one: no other group runs beside it on a
    // concurrent graph branch
    bool solo = false;
};

struct TuneResult {
    Tune sched;
    bool tuned = false;       // measured (else the process-wide schedule)
    float us_default = 0.f;   // subset time with the default schedule
    float us_best = 0.f;      // subset time with `sched`
    bool cached = false;      // read from the MINFFT_TUNE_CACHE file, not measured here
};

bool autotune_enabled();
// measures candidate schedules for the group on `st` (synchronous); falls
// back to the process-wide schedule when the group is small or scratch
// memory is short
TuneResult autotune_group(const Workspace& ws, const TuneRequest& r, hipStream_t st);
std::string describe_tune(const Tune& t);

}  // namespace mfx
