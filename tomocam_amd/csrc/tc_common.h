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

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    TC_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
  }
  // grows (never shrinks) the allocation
  void reserve(size_t bytes) {
    if (bytes > n_) alloc(bytes);
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p_); }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
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
  KC_COUNT = 6,
};

// Begins/ends a timed region for kernel class `kc` on `stream` if timing for
// that class is enabled; cheap no-op otherwise.
// `bytes` is the algorithmic HBM traffic of the launch (inputs read once,
// outputs written once), accumulated for the roofline report.
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
