// Live per-kernel timing with HIP events on the launching stream.
//
// One kernel, selected by name, is timed: every launch of it is bracketed by a
// pair of events recorded on the stream it runs on, and the launcher adds the
// algorithmic bytes it moves. ktimer_read() sums the event intervals. Disabled,
// a KScope costs one branch.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace impes {

void ktimer_config(const char* kernel, bool enable);  // resets the totals
void ktimer_read(std::string* kernel, double* ms, double* bytes, long long* launches);
bool ktimer_enabled();  // graph replays are bypassed while a kernel is being timed

class KScope {
 public:
  KScope(const char* kernel, double bytes, hipStream_t s);
  ~KScope();
  KScope(const KScope&) = delete;
  KScope& operator=(const KScope&) = delete;

 private:
  int slot_ = -1;
  hipStream_t s_ = nullptr;
};

}  // namespace impes
