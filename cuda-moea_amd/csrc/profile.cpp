// Opt-in per-stage timing with HIP events on the launch stream (the role of
// the reference's ASAP_NDSORT_FRONT_PROP_TIMING breakdown,
// non_dominated_sort.cu:101-107, 1132-1276).  Off by default; when enabled,
// each instrumented stage records an event pair on its stream and
// moea_profile_read sums the elapsed times.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

const char* kStageNames[kNumStages] = {
    "reproduce", "evaluate", "nds_dominance", "nds_peel", "select_prep",
    "association", "niching", "gather", "rvea_select", "rvea_adapt", "nds_sort",
};

struct Timer {
    std::mutex mu;
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<hipEvent_t> free;             // events not held by an open or unread record
    struct Rec { int stage; hipEvent_t a, b; };
    std::vector<Rec> open;
    std::vector<Rec> done;
    double ms[kNumStages] = {};
    long long count[kNumStages] = {};
};

Timer& timer()
{
    static Timer t;
    return t;
}

hipEvent_t take(Timer& t)
{
    if (t.free.empty()) return nullptr;
    hipEvent_t e = t.free.back();
    t.free.pop_back();
    return e;
}

void drain(Timer& t)
{
    for (auto& r : t.done) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            t.ms[r.stage] += ms;
            t.count[r.stage] += 1;
        }
        t.free.push_back(r.a);
        t.free.push_back(r.b);
    }
    t.done.clear();
}

// Launches being captured into a graph are not timed: an event recorded
// during capture would become a graph node.
bool capturing(hipStream_t stream)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

}  // namespace

void stage_begin(int stage, hipStream_t stream)
{
    Timer& t = timer();
    if (!t.enabled || capturing(stream)) return;
    std::lock_guard<std::mutex> lk(t.mu);
    if (t.free.size() < 2) drain(t);
    if (t.free.size() < 2) return;
    hipEvent_t a = take(t), b = take(t);
    (void)hipEventRecord(a, stream);
    t.open.push_back({stage, a, b});
}

void stage_end(int stage, hipStream_t stream)
{
    Timer& t = timer();
    if (!t.enabled || capturing(stream)) return;
    std::lock_guard<std::mutex> lk(t.mu);
    for (size_t i = t.open.size(); i-- > 0;) {
        if (t.open[i].stage == stage) {
            (void)hipEventRecord(t.open[i].b, stream);
            t.done.push_back(t.open[i]);
            t.open.erase(t.open.begin() + static_cast<long>(i));
            return;
        }
    }
}

}  // namespace moea

using namespace moea;

extern "C" {

int moea_profile_enable(int enable)
{
    Timer& t = timer();
    std::lock_guard<std::mutex> lk(t.mu);
    if (enable && !t.enabled) {
        t.pool.resize(8192);
        for (auto& e : t.pool) MOEA_HIP(hipEventCreate(&e));
        t.free = t.pool;
        for (int i = 0; i < kNumStages; ++i) { t.ms[i] = 0.0; t.count[i] = 0; }
        t.enabled = true;
    } else if (!enable && t.enabled) {
        drain(t);
        for (auto& e : t.pool) hipEventDestroy(e);
        t.pool.clear();
        t.free.clear();
        t.open.clear();
        t.enabled = false;
    }
    return kOk;
}

int moea_profile_read(double* ms_out, long long* count_out, int max_stages)
{
    Timer& t = timer();
    std::lock_guard<std::mutex> lk(t.mu);
    drain(t);
    const int n = max_stages < kNumStages ? max_stages : kNumStages;
    for (int i = 0; i < n; ++i) {
        if (ms_out) ms_out[i] = t.ms[i];
        if (count_out) count_out[i] = t.count[i];
        t.ms[i] = 0.0;
        t.count[i] = 0;
    }
    return kNumStages;
}

const char* moea_profile_stage_name(int stage)
{
    return (stage >= 0 && stage < kNumStages) ? kStageNames[stage] : "";
}

}  // extern "C"
