// Live per-kernel timing with HIP events on the launching stream.
//
// One kernel, selected by name, is timed. A KScope naming it arms the timer for
// the next klaunch() on this thread: that launch goes through hipExtLaunchKernel
// with the scope's start/stop events, so the dispatch itself stamps them at the
// kernel's start and end (the interval a profiler reports; no separate event
// packets and their dispatch latency inside it). The launcher adds the
// algorithmic bytes it moves. ktimer_read() sums the intervals. Disabled, a
// KScope costs one branch and klaunch is a plain launch.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <string>

namespace impes {

void ktimer_config(const char* kernel, bool enable);  // resets the totals
void ktimer_read(std::string* kernel, double* ms, double* bytes, long long* launches);
bool ktimer_enabled();  // graph replays are bypassed while a kernel is being timed

struct KSlot {
  hipEvent_t start = nullptr, stop = nullptr;
};
KSlot* ktimer_armed();  // this thread's armed slot, or nullptr
void ktimer_consume();  // the armed slot was handed to a launch

class KScope {
 public:
  KScope(const char* kernel, double bytes, hipStream_t s);
  ~KScope();
  KScope(const KScope&) = delete;
  KScope& operator=(const KScope&) = delete;

 private:
  int slot_ = -1;
  double bytes_ = 0.0;
  hipStream_t s_ = nullptr;
};

template <typename F, typename... Args>
inline void klaunch(F kernel, dim3 grid, dim3 block, unsigned shmem, hipStream_t s, Args... args) {
  if (KSlot* k = ktimer_armed()) {
    ktimer_consume();
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, k->start, k->stop, 0u, args...);
  } else {
    kernel<<<grid, block, shmem, s>>>(args...);
  }
}

}  // namespace impes
