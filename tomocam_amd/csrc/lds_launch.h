// Launch-shape selection for kernels that keep whole FFT lines in LDS.
#pragma once

#include "fft_device.h"

namespace tc {

struct LdsLaunch {
  int lpw = 1;     // lines per workgroup
  int nthr = 256;  // threads per workgroup
  int nb = 1;      // butterflies per thread per stage (template parameter)
  size_t smem = 0; // dynamic LDS bytes
};

// Picks lines-per-workgroup <= lpw_want whose LDS footprint (plus
// `extra_bytes_per_line`) stays under `smem_cap`, and the smallest thread
// count that needs at most 2 (else 4) butterflies per thread per stage.
inline LdsLaunch plan_lds_launch(const FFTLength& F, int lpw_want,
                                 size_t smem_cap,
                                 size_t extra_bytes_per_line = 0) {
  LdsLaunch r;
  const size_t per_line = sizeof(float2) * (size_t)F.L() + extra_bytes_per_line;
  int lpw = std::max(1, std::min<int>(lpw_want, (int)(smem_cap / per_line)));
  for (;;) {
    for (int limit : {2, 4}) {
      for (int nthr : {256, 512, 1024}) {
        int nb = F.nb_for(lpw, nthr);
        if (nb != 0 && nb <= limit) {
          r.lpw = lpw;
          r.nthr = nthr;
          r.nb = nb;
          r.smem = per_line * lpw;
          TC_REQUIRE(r.smem <= 160 * 1024, "FFT line does not fit in LDS");
          return r;
        }
      }
    }
    TC_REQUIRE(lpw > 1, "FFT length too large for one workgroup");
    lpw /= 2;
  }
}

// Raises the dynamic-LDS limit of a kernel once per process.
inline void allow_lds(const void* kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    TC_HIP(hipFuncSetAttribute(kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

// Dispatch helper: calls f.template operator()<NB>() for NB in {1,2,4}.
#define TC_DISPATCH_NB(nbval, ...)                          \
  switch (nbval) {                                          \
    case 1: { constexpr int NB = 1; __VA_ARGS__; } break;   \
    case 2: { constexpr int NB = 2; __VA_ARGS__; } break;   \
    case 4: { constexpr int NB = 4; __VA_ARGS__; } break;   \
    default:                                                \
      throw ::tc::Error(TC_ERR_INTERNAL, "bad NB");   \
  }

// Block-wide sum of a double; returns the total in thread 0.  Uses the
// caller-provided LDS array of at least 16 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) t += red[i];
  }
  __syncthreads();
  return t;
}

}  // namespace tc
