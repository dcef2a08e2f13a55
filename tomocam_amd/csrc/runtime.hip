// Error strings and optional per-kernel-class event timing.
#include <mutex>
#include <vector>

#include "tc_common.h"

namespace tc {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// Event pairs recorded around launches of enabled kernel classes; resolved
// (and released) by ktimer_collect.
struct KTimer {
  std::mutex mu;
  bool enabled[KC_COUNT] = {};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs[KC_COUNT];
  double bytes[KC_COUNT] = {};
};

// the event opened by ktimer_begin, per thread (concurrent callers)
static thread_local hipEvent_t t_open[KC_COUNT] = {};

static KTimer& kt() {
  static KTimer k;
  return k;
}

void ktimer_enable(int kc, bool on) {
  TC_REQUIRE(kc >= 0 && kc < KC_COUNT, "bad kernel class");
  std::lock_guard<std::mutex> lk(kt().mu);
  kt().enabled[kc] = on;
}

bool ktimer_active() {
  KTimer& k = kt();
  std::lock_guard<std::mutex> lk(k.mu);
  for (bool on : k.enabled)
    if (on) return true;
  return false;
}

void ktimer_begin(int kc, hipStream_t stream) {
  KTimer& k = kt();
  if (!k.enabled[kc]) return;
  hipEvent_t e;
  TC_HIP(hipEventCreate(&e));
  TC_HIP(hipEventRecord(e, stream));
  t_open[kc] = e;
}

void ktimer_end(int kc, hipStream_t stream, double bytes) {
  KTimer& k = kt();
  if (!k.enabled[kc] || !t_open[kc]) return;
  hipEvent_t e;
  TC_HIP(hipEventCreate(&e));
  TC_HIP(hipEventRecord(e, stream));
  std::lock_guard<std::mutex> lk(k.mu);
  k.pairs[kc].emplace_back(t_open[kc], e);
  k.bytes[kc] += bytes;
  t_open[kc] = nullptr;
}

// Total milliseconds, launch count and algorithmic bytes of class kc since
// the last collect.
void ktimer_collect(int kc, double* total_ms, long* count, double* bytes) {
  TC_REQUIRE(kc >= 0 && kc < KC_COUNT, "bad kernel class");
  KTimer& k = kt();
  std::lock_guard<std::mutex> lk(k.mu);
  double tot = 0.0;
  long cnt = 0;
  for (auto& pr : k.pairs[kc]) {
    TC_HIP(hipEventSynchronize(pr.second));
    float ms = 0.f;
    TC_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    tot += ms;
    ++cnt;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  k.pairs[kc].clear();
  *total_ms = tot;
  *count = cnt;
  *bytes = k.bytes[kc];
  k.bytes[kc] = 0.0;
}

}  // namespace tc
