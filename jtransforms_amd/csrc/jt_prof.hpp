// Optional per-launch timing with HIP events, recorded on the stream each
// kernel is launched on (used by the benchmark to measure kernel durations
// live; off by default, no cost when off).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace jt {

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};

inline std::atomic<bool>& prof_on() {
    static std::atomic<bool> on{false};
    return on;
}
inline std::mutex& prof_mu() {
    static std::mutex m;
    return m;
}
inline std::vector<ProfRec>& prof_list() {
    static std::vector<ProfRec> v;
    return v;
}

// JT_LAUNCH_LOG=<path>: every launch label is appended to that file in
// launch order (one line per kernel dispatch), so that a profiler's
// per-dispatch records can be matched to the library's launches.
inline FILE* launch_log() {
    static FILE* f = []() -> FILE* {
        const char* p = std::getenv("JT_LAUNCH_LOG");
        return (p && p[0]) ? std::fopen(p, "a") : nullptr;
    }();
    return f;
}

// Records an event pair around the launches made in its lifetime.
class ProfScope {
   public:
    template <class NameFn>
    ProfScope(hipStream_t st, NameFn&& fn) : st_(st) {
        if (FILE* lf = launch_log()) {
            const std::string label = fn();
            std::lock_guard<std::mutex> g(prof_mu());
            std::fprintf(lf, "%s\n", label.c_str());
            std::fflush(lf);
        }
        if (!prof_on().load(std::memory_order_relaxed)) return;
        // launches recorded into a graph under stream capture are not timed
        // (their events would only be recorded when the graph replays)
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
        active_ = true;
        name_ = fn();
        // timing-only events: no system-scope fence (cache writeback and
        // invalidate) when recorded, so the measured kernels run as they
        // would unprofiled
        (void)hipEventCreateWithFlags(&a_, hipEventDisableSystemFence);
        (void)hipEventCreateWithFlags(&b_, hipEventDisableSystemFence);
        (void)hipEventRecord(a_, st_);
    }
    ~ProfScope() {
        if (!active_) return;
        (void)hipEventRecord(b_, st_);
        std::lock_guard<std::mutex> g(prof_mu());
        prof_list().push_back(ProfRec{name_, a_, b_});
    }

   private:
    hipStream_t st_;
    bool active_ = false;
    std::string name_;
    hipEvent_t a_{}, b_{};
};

inline void prof_clear() {
    std::lock_guard<std::mutex> g(prof_mu());
    for (auto& r : prof_list()) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    prof_list().clear();
}

// "name<TAB>count<TAB>total_ms" lines, sorted by name.
inline std::string prof_summary() {
    std::lock_guard<std::mutex> g(prof_mu());
    std::map<std::string, std::pair<long, double>> agg;
    for (auto& r : prof_list()) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        auto& e = agg[r.name];
        e.first += 1;
        e.second += ms;
    }
    std::string out;
    char line[512];
    for (auto& kv : agg) {
        snprintf(line, sizeof line, "%s\t%ld\t%.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    return out;
}

}  // namespace jt

#define JT_PROF(st, ...) ::jt::ProfScope jt_prof_scope_##__LINE__((st), [&]() { return std::string(__VA_ARGS__); })
