// Optional per-kernel-class launch timing with HIP events, for measurement only.
//
// When enabled, the launch sites of the heavy kernels bracket each launch with a pair of
// events recorded on the launch stream and add the launch's algorithmic work (flops or
// pair-features) to the class total. qmlh_timer_read synchronises on the recorded events and
// reports total time, launch count and work per class. Disabled (the default), a launch site
// costs one relaxed atomic load.
#include "timer.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace qmlh {

namespace {
struct Pending {
    hipEvent_t start, stop;
    int cls;
    double work;
};
std::atomic<int> g_enabled{0};
std::mutex g_mu;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
double g_ms[QMLH_TIMER_CLASSES];
double g_work[QMLH_TIMER_CLASSES];
int64_t g_count[QMLH_TIMER_CLASSES];

hipEvent_t take_event() {
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain_locked() {
    for (auto& p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            g_ms[p.cls] += ms;
            g_work[p.cls] += p.work;
            g_count[p.cls] += 1;
        }
        g_free.push_back(p.start);
        g_free.push_back(p.stop);
    }
    g_pending.clear();
}
}  // namespace

bool timer_on() { return g_enabled.load(std::memory_order_relaxed) != 0; }

TimerScope::TimerScope(int cls, double work, hipStream_t s) : cls_(cls), work_(work), s_(s) {
    if (!timer_on() || cls < 0 || cls >= QMLH_TIMER_CLASSES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    start_ = take_event();
    stop_ = take_event();
    if (start_ && stop_ && hipEventRecord(start_, s_) == hipSuccess) active_ = true;
}

TimerScope::~TimerScope() {
    if (!active_) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (hipEventRecord(stop_, s_) == hipSuccess) {
        g_pending.push_back({start_, stop_, cls_, work_});
    } else {
        g_free.push_back(start_);
        g_free.push_back(stop_);
    }
}

}  // namespace qmlh

extern "C" int qmlh_timer_enable(int32_t on) {
    using namespace qmlh;
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    if (on) {
        for (int c = 0; c < QMLH_TIMER_CLASSES; ++c) g_ms[c] = g_work[c] = 0.0, g_count[c] = 0;
    }
    g_enabled.store(on ? 1 : 0);
    return QMLH_OK;
}

extern "C" int qmlh_timer_read(int32_t cls, double* total_ms_h, int64_t* launches_h,
                               double* work_h) {
    using namespace qmlh;
    if (cls < 0 || cls >= QMLH_TIMER_CLASSES)
        return set_error(QMLH_ERR_ARG, "timer_read: class %d out of range", cls);
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    if (total_ms_h) *total_ms_h = g_ms[cls];
    if (launches_h) *launches_h = g_count[cls];
    if (work_h) *work_h = g_work[cls];
    return QMLH_OK;
}
