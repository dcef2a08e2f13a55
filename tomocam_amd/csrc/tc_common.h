// Common host/device helpers for the tomoCAM MI355X hot path (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include "tomocam_hip.h"  // status codes TC_OK, TC_ERR_*

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tc {

// ---------------------------------------------------------------------------
// errors: C-ABI entry points catch these and turn them into status codes
// ---------------------------------------------------------------------------

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

#define TC_HIP(call)                                                        \
  do {                                                                      \
    hipError_t tc_e_ = (call);                                              \
    if (tc_e_ != hipSuccess) {                                              \
      throw ::tc::Error(tc_e_ == hipErrorOutOfMemory ? TC_ERR_NOMEM   \
                                                     : TC_ERR_HIP,    \
                        std::string(#call) + ": " +                         \
                            hipGetErrorString(tc_e_) + " (" __FILE__ ":" +  \
                            std::to_string(__LINE__) + ")");                \
    }                                                                       \
  } while (0)

#define TC_REQUIRE(cond, msg)                                              \
  do {                                                                     \
    if (!(cond)) throw ::tc::Error(TC_ERR_INVALID, (msg));           \
  } while (0)

// Runs a C-ABI body, mapping exceptions to status codes.
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return TC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return TC_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return TC_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

// Device-memory cache of the calling host thread.  Inside an ArenaScope,
// DevBuf takes its memory from the thread's arena: released buffers go back
// to the arena's free list (not to hipFree, which synchronises the whole
// device) and later requests of a similar size reuse them, so a repeated
// MBIR call allocates nothing after the first.  An arena is owned by one
// host thread at a time; when the thread ends, its arena goes back to a
// process-wide pool for the next thread, with its blocks still cached.
// Reuse is ordered by stream work:
//   - ArenaScope() belongs around work that synchronises its streams
//     before the scope ends (the MBIR drivers);
//   - ArenaScope(stream) around work left pending on `stream` (the
//     asynchronous operators): the outermost such scope records an event
//     on `stream` when it ends, and a later scope on another stream (or a
//     synchronising one) waits for that event before any block is reused.
struct Arena;
class ArenaScope {
 public:
  ArenaScope();
  explicit ArenaScope(hipStream_t stream);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;

 private:
  void enter(bool async, hipStream_t stream);
  bool async_ = false;
  hipStream_t stream_ = nullptr;
};
// Autotuning on the calling thread: on unless TOMOCAM_HIP_AUTOTUNE=0 or an
// AutotuneScope(false) is active (concurrent batch workers, whose timings
// would include the other workers' kernels).
bool autotune_enabled();
class AutotuneScope {
 public:
  explicit AutotuneScope(bool on);
  ~AutotuneScope();
  AutotuneScope(const AutotuneScope&) = delete;
  AutotuneScope& operator=(const AutotuneScope&) = delete;

 private:
  bool saved_;
};
// TOMOCAM_HIP_TUNE_LOG=1: the autotuners print each candidate's time to
// stderr.
bool tune_log();

// Turns the calling thread's arena off for its lifetime: long-lived objects
// (FFT twiddle tables, NUFFT and ray-DFT plans) allocate fresh memory,
// because they fill it with synchronous copies on the null stream, which
// is not ordered after the pending work of the thread's stream that a
// recycled block may still belong to.
class ArenaPause {
 public:
  ArenaPause();
  ~ArenaPause();
  ArenaPause(const ArenaPause&) = delete;
  ArenaPause& operator=(const ArenaPause&) = delete;

 private:
  int saved_;
};
// A cached block of at least `bytes` from the thread's arena when a scope is
// active (its arena and capacity returned), else nullptr.
void* arena_alloc(size_t bytes, Arena** owner, size_t* cap);
// Returns a block to `owner` (its owning thread) or frees it.
void arena_free(void* p, size_t cap, Arena* owner);
// Device allocations made by arenas, requests served from a cache, and the
// bytes cached (free) across all arenas.
void arena_stats(long* mallocs, long* hits, double* cached_bytes);
// Frees every cached (free) block of every pooled arena and of the calling
// thread's arena.
void arena_trim();
// Eight doubles of pinned host memory belonging to the calling thread's
// arena (inside an ArenaScope), kept with the arena for the process.
double* arena_pinned_dots();

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept
      : p_(o.p_), n_(o.n_), cap_(o.cap_), arena_(o.arena_) {
    o.p_ = nullptr;
    o.n_ = o.cap_ = 0;
    o.arena_ = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      cap_ = o.cap_;
      arena_ = o.arena_;
      o.p_ = nullptr;
      o.n_ = o.cap_ = 0;
      o.arena_ = nullptr;
    }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    p_ = arena_alloc(bytes, &arena_, &cap_);
    if (!p_) {
      TC_HIP(hipMalloc(&p_, bytes));
      cap_ = bytes;
    }
    n_ = bytes;
  }
  // grows (never shrinks) the allocation
  void reserve(size_t bytes) {
    if (bytes <= n_) return;
    if (p_ && bytes <= cap_) {
      n_ = bytes;
      return;
    }
    alloc(bytes);
  }
  void release() {
    if (p_) {
      if (arena_)
        arena_free(p_, cap_, arena_);
      else
        (void)hipFree(p_);
    }
    p_ = nullptr;
    n_ = cap_ = 0;
    arena_ = nullptr;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p_); }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
  size_t cap_ = 0;
  Arena* arena_ = nullptr;
};

inline unsigned cdiv(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// kernel timing hooks (used by the benchmark to time the dominant kernel)
// ---------------------------------------------------------------------------

enum KernelClass : int {
  KC_TOEPLITZ_ROWS_FWD = 0,
  KC_TOEPLITZ_COLS = 1,
  KC_TOEPLITZ_ROWS_INV = 2,
  KC_MBIR_GRADIENT = 3,
  KC_MBIR_COMMIT = 4,
  KC_QGGMRF = 5,
  // the FP16-storage Toeplitz passes (mixed precision), timed apart
  KC_TOEPLITZ_ROWS_FWD_F16 = 6,
  KC_TOEPLITZ_COLS_F16 = 7,
  KC_TOEPLITZ_ROWS_INV_F16 = 8,
  KC_COUNT = 9,
};

// Begins/ends a timed region for kernel class `kc` on `stream` if timing for
// that class is enabled; cheap no-op otherwise.
// `bytes` is the algorithmic HBM traffic of the launch (inputs read once,
// outputs written once), accumulated for the roofline report.
void ktimer_set_stride(int stride);
void ktimer_begin(int kc, hipStream_t stream);
void ktimer_end(int kc, hipStream_t stream, double bytes = 0.0);
// True if timing is enabled for any kernel class.
bool ktimer_active();

// ---------------------------------------------------------------------------
// device complex helpers (float2 = re, im)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

}  // namespace tc

namespace tc {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup i runs on XCD i mod 8).  This maps the hardware index to a
// logical block index so that logical neighbours (which touch neighbouring
// memory) land on the same XCD and share its L2.  Bijective on [0, n).
__device__ __forceinline__ int xcd_remap(int i, int n) {
  const int q = n >> 3;
  const int rmd = n & 7;
  const int x = i & 7;
  const int t = i >> 3;
  const int start = x * q + (x < rmd ? x : rmd);
  return start + t;
}

}  // namespace tc
