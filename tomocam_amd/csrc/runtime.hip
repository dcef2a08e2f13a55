// Error strings, the per-thread device-memory arenas and optional
// per-kernel-class event timing.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "tc_common.h"

namespace tc {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// arenas
// ---------------------------------------------------------------------------

struct Arena {
  std::mutex mu;
  std::multimap<size_t, void*> free;  // capacity -> block
  size_t free_bytes = 0;
  int dev = -1;
  double* pinned = nullptr;  // arena_pinned_dots
  // the last asynchronous scope's stream and its completion event
  hipStream_t last = nullptr;
  hipEvent_t done = nullptr;
  bool pending = false;
};

static std::mutex g_pool_mu;
static std::vector<Arena*> g_pool;        // arenas no thread owns
static std::vector<Arena*> g_all;         // every arena (never destroyed)
static std::atomic<long> g_mallocs{0}, g_hits{0};
static std::atomic<long long> g_free_bytes{0};

// the calling thread's arena; back to the pool when the thread ends
struct ArenaOwner {
  Arena* a = nullptr;
  int depth = 0;
  ~ArenaOwner() {
    if (!a) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(a);
  }
};
static thread_local ArenaOwner t_arena;

ArenaScope::ArenaScope() { enter(false, nullptr); }
ArenaScope::ArenaScope(hipStream_t stream) { enter(true, stream); }

void ArenaScope::enter(bool async, hipStream_t stream) {
  async_ = async;
  stream_ = stream;
  if (!t_arena.a) {
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = std::find_if(g_pool.begin(), g_pool.end(),
                           [&](Arena* a) { return a->dev == dev; });
    if (it != g_pool.end()) {
      t_arena.a = *it;
      g_pool.erase(it);
    } else {
      t_arena.a = new Arena;
      t_arena.a->dev = dev;
      g_all.push_back(t_arena.a);
    }
  }
  // blocks released by an asynchronous scope may still be in use on its
  // stream: wait unless this scope runs on that same stream
  Arena* a = t_arena.a;
  if (a->pending && (!async || stream != a->last)) {
    TC_HIP(hipEventSynchronize(a->done));
    a->pending = false;
  }
  ++t_arena.depth;
}

ArenaScope::~ArenaScope() {
  if (--t_arena.depth != 0 || !async_) return;
  Arena* a = t_arena.a;
  if (!a->done &&
      hipEventCreateWithFlags(&a->done, hipEventDisableTiming) != hipSuccess) {
    // no event: fall back to waiting for the stream now
    (void)hipGetLastError();
    (void)hipStreamSynchronize(stream_);
    return;
  }
  if (hipEventRecord(a->done, stream_) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(stream_);
    return;
  }
  a->last = stream_;
  a->pending = true;
}

static thread_local bool t_autotune = true;

bool autotune_enabled() {
  const char* e = std::getenv("TOMOCAM_HIP_AUTOTUNE");
  return t_autotune && !(e && e[0] == '0');
}

AutotuneScope::AutotuneScope(bool on) : saved_(t_autotune) {
  t_autotune = t_autotune && on;
}
AutotuneScope::~AutotuneScope() { t_autotune = saved_; }

bool tune_log() {
  const char* e = std::getenv("TOMOCAM_HIP_TUNE_LOG");
  return e && e[0] == '1';
}

ArenaPause::ArenaPause() : saved_(t_arena.depth) { t_arena.depth = 0; }
ArenaPause::~ArenaPause() { t_arena.depth = saved_; }

static constexpr size_t kArenaGrain = size_t(2) << 20;  // 2 MiB

// frees cached blocks of `a` (caller holds a->mu), largest first, until at
// most `keep` bytes stay cached
static void arena_shrink(Arena* a, size_t keep) {
  while (a->free_bytes > keep && !a->free.empty()) {
    auto it = std::prev(a->free.end());
    (void)hipFree(it->second);
    a->free_bytes -= it->first;
    g_free_bytes -= (long long)it->first;
    a->free.erase(it);
  }
}

// cached bytes beyond this share of the device's memory are freed
static size_t cache_limit() {
  static size_t lim = [] {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || tot == 0)
      return size_t(16) << 30;
    return tot / 5 * 2;
  }();
  return lim;
}

void* arena_alloc(size_t bytes, Arena** owner, size_t* cap) {
  Arena* a = t_arena.a;
  if (!a || t_arena.depth == 0) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != a->dev) return nullptr;
  const size_t want = (bytes + kArenaGrain - 1) / kArenaGrain * kArenaGrain;
  {
    std::lock_guard<std::mutex> lk(a->mu);
    // the smallest cached block that fits, if it wastes at most half
    auto it = a->free.lower_bound(want);
    if (it != a->free.end() && it->first <= 2 * want) {
      void* p = it->second;
      *cap = it->first;
      a->free_bytes -= it->first;
      g_free_bytes -= (long long)it->first;
      a->free.erase(it);
      *owner = a;
      ++g_hits;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    // out of memory: give the cache back and try once more
    {
      std::lock_guard<std::mutex> lk(a->mu);
      arena_shrink(a, 0);
    }
    arena_trim();
    TC_HIP(hipMalloc(&p, want));
  }
  ++g_mallocs;
  *cap = want;
  *owner = a;
  return p;
}

void arena_free(void* p, size_t cap, Arena* owner) {
  if (owner != t_arena.a) {
    // another thread's block: its stream work may still be pending, and
    // hipFree waits for the device
    (void)hipFree(p);
    return;
  }
  std::lock_guard<std::mutex> lk(owner->mu);
  owner->free.emplace(cap, p);
  owner->free_bytes += cap;
  g_free_bytes += (long long)cap;
  if ((size_t)std::max<long long>(0, g_free_bytes.load()) > cache_limit())
    arena_shrink(owner, owner->free_bytes / 2);
}

double* arena_pinned_dots() {
  Arena* a = t_arena.a;
  TC_REQUIRE(a && t_arena.depth > 0, "pinned dots need an arena scope");
  if (!a->pinned)
    TC_HIP(hipHostMalloc((void**)&a->pinned, sizeof(double) * 8,
                         hipHostMallocDefault));
  return a->pinned;
}

void arena_stats(long* mallocs, long* hits, double* cached_bytes) {
  if (mallocs) *mallocs = g_mallocs.load();
  if (hits) *hits = g_hits.load();
  if (cached_bytes) *cached_bytes = (double)g_free_bytes.load();
}

void arena_trim() {
  std::vector<Arena*> as;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    as = g_pool;
  }
  if (t_arena.a) as.push_back(t_arena.a);
  for (Arena* a : as) {
    std::lock_guard<std::mutex> lk(a->mu);
    arena_shrink(a, 0);
  }
}

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------

// Event pairs recorded around launches of enabled kernel classes; resolved
// by ktimer_collect, which returns the events to a per-device pool (creating
// two events per timed launch measurably slowed loops of short launches).
struct KPair {
  hipEvent_t a, b;
  int dev;
};

struct KTimer {
  std::mutex mu;
  bool enabled[KC_COUNT] = {};
  std::vector<KPair> pairs[KC_COUNT];
  double bytes[KC_COUNT] = {};
  std::map<int, std::vector<hipEvent_t>> pool;  // resolved events, by device
  // one launch in `stride` of each class is timed (launch counters per class)
  std::atomic<int> stride{1};
  std::atomic<long> launches[KC_COUNT] = {};
};

// the event opened by ktimer_begin (and its device), per thread (concurrent
// callers)
static thread_local hipEvent_t t_open[KC_COUNT] = {};
static thread_local int t_open_dev[KC_COUNT] = {};

static KTimer& kt() {
  static KTimer k;
  return k;
}

// an event of device dev from the pool, else a new one
static hipEvent_t ktimer_event(KTimer& k, int dev) {
  {
    std::lock_guard<std::mutex> lk(k.mu);
    auto& v = k.pool[dev];
    if (!v.empty()) {
      hipEvent_t e = v.back();
      v.pop_back();
      return e;
    }
  }
  hipEvent_t e;
  TC_HIP(hipEventCreate(&e));
  return e;
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

void ktimer_set_stride(int stride) {
  TC_REQUIRE(stride >= 1, "kernel timer: stride must be >= 1");
  KTimer& k = kt();
  k.stride.store(stride);
  for (auto& c : k.launches) c.store(0);
}

void ktimer_begin(int kc, hipStream_t stream) {
  KTimer& k = kt();
  if (!k.enabled[kc]) return;
  const int st = k.stride.load();
  if (st > 1 && k.launches[kc].fetch_add(1) % st != 0) return;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  hipEvent_t e = ktimer_event(k, dev);
  TC_HIP(hipEventRecord(e, stream));
  t_open[kc] = e;
  t_open_dev[kc] = dev;
}

void ktimer_end(int kc, hipStream_t stream, double bytes) {
  KTimer& k = kt();
  if (!k.enabled[kc] || !t_open[kc]) return;
  const int dev = t_open_dev[kc];
  hipEvent_t e = ktimer_event(k, dev);
  TC_HIP(hipEventRecord(e, stream));
  std::lock_guard<std::mutex> lk(k.mu);
  k.pairs[kc].push_back(KPair{t_open[kc], e, dev});
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
    TC_HIP(hipEventSynchronize(pr.b));
    float ms = 0.f;
    TC_HIP(hipEventElapsedTime(&ms, pr.a, pr.b));
    tot += ms;
    ++cnt;
    auto& v = k.pool[pr.dev];
    v.push_back(pr.a);
    v.push_back(pr.b);
  }
  k.pairs[kc].clear();
  *total_ms = tot;
  *count = cnt;
  *bytes = k.bytes[kc];
  k.bytes[kc] = 0.0;
}

}  // namespace tc
