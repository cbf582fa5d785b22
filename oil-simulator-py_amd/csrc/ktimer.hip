// Event-pair timer of one selected kernel (see ktimer.h).
#include "ktimer.h"

#include <cstring>
#include <vector>

#include "common.h"

namespace impes {

namespace {

constexpr int kPairs = 2048;  // event pairs in flight before the totals are folded

struct State {
  bool on = false;
  std::string target;
  std::vector<hipEvent_t> ev;  // start = ev[2k], end = ev[2k+1]
  int used = 0;
  double ms = 0.0, bytes = 0.0;
  long long launches = 0;
};

State& st() {
  static State s;
  return s;
}

void fold() {
  State& S = st();
  for (int k = 0; k < S.used; ++k) {
    IMPES_HIP_CHECK(hipEventSynchronize(S.ev[2 * k + 1]));
    float ms = 0.f;
    IMPES_HIP_CHECK(hipEventElapsedTime(&ms, S.ev[2 * k], S.ev[2 * k + 1]));
    S.ms += ms;
  }
  S.used = 0;
}

}  // namespace

void ktimer_config(const char* kernel, bool enable) {
  State& S = st();
  if (S.used) fold();
  S.on = enable && kernel && kernel[0];
  S.target = S.on ? kernel : "";
  S.ms = S.bytes = 0.0;
  S.launches = 0;
  if (S.on && S.ev.empty()) {
    S.ev.resize(2 * kPairs);
    for (auto& e : S.ev) IMPES_HIP_CHECK(hipEventCreate(&e));
  }
}

void ktimer_read(std::string* kernel, double* ms, double* bytes, long long* launches) {
  State& S = st();
  fold();
  if (kernel) *kernel = S.target;
  if (ms) *ms = S.ms;
  if (bytes) *bytes = S.bytes;
  if (launches) *launches = S.launches;
}

bool ktimer_enabled() { return st().on; }

namespace {
thread_local KSlot t_armed;
thread_local bool t_is_armed = false;
}  // namespace

KSlot* ktimer_armed() { return t_is_armed ? &t_armed : nullptr; }
void ktimer_consume() { t_is_armed = false; }

KScope::KScope(const char* kernel, double bytes, hipStream_t s) {
  State& S = st();
  if (!S.on || S.target != kernel) return;
  if (S.used == kPairs) fold();
  slot_ = S.used++;
  bytes_ = bytes;
  s_ = s;
  S.bytes += bytes;
  S.launches += 1;
  t_armed = KSlot{S.ev[2 * slot_], S.ev[2 * slot_ + 1]};
  t_is_armed = true;
}

KScope::~KScope() {
  if (slot_ < 0) return;
  if (t_is_armed) {  // nothing was launched (empty operand): an empty interval, counts taken back
    t_is_armed = false;
    State& S = st();
    S.bytes -= bytes_;
    S.launches -= 1;
    (void)hipEventRecord(S.ev[2 * slot_], s_);
    (void)hipEventRecord(S.ev[2 * slot_ + 1], s_);
  }
}

}  // namespace impes
