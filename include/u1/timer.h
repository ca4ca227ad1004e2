// timer.h -- drop-in for the reference's wall-clock Timer (timer.cpp:61-105).
#ifndef U1_DROPIN_TIMER_H
#define U1_DROPIN_TIMER_H
#include <chrono>

namespace U1 {
class Timer {
 public:
  void start() { t0_ = clock::now(); running_ = true; }
  void stop() { t1_ = clock::now(); running_ = false; }
  double getElapsedTime() const { return getElapsedTimeInSec(); }
  double getElapsedTimeInSec() const {
    const auto t1 = running_ ? clock::now() : t1_;
    return std::chrono::duration<double>(t1 - t0_).count();
  }
  double getElapsedTimeInMilliSec() const { return 1e3 * getElapsedTimeInSec(); }

 private:
  using clock = std::chrono::steady_clock;
  clock::time_point t0_{clock::now()}, t1_{t0_};
  bool running_ = false;
};
}  // namespace U1
#endif
