// Launch timing hooks (see timer.hip).
#pragma once

#include "common.h"

namespace qmlh {

bool timer_on();

// Records an event pair around its lifetime on stream `s` when timing is enabled.
class TimerScope {
   public:
    TimerScope(int cls, double work, hipStream_t s);
    ~TimerScope();
    // Work known only at the end (iterations, for instance): replaces the constructor's.
    void set_work(double work) { work_ = work; }
    TimerScope(const TimerScope&) = delete;
    TimerScope& operator=(const TimerScope&) = delete;

   private:
    int cls_;
    double work_;
    hipStream_t s_;
    hipEvent_t start_ = nullptr, stop_ = nullptr;
    bool active_ = false;
};

}  // namespace qmlh
