// One radix-16 DFT stage of a Stockham FFT on LDS-resident data, computed
// two ways on gfx950, to measure whether a DFT-as-GEMM stage on the matrix
// cores beats the VALU butterflies the Toeplitz passes use:
//
//   valu: each thread holds one 16-point butterfly and computes it as 4 x 4
//         (radix-4, twiddles, radix-4) in FP32 registers;
//   mfma: each wave computes 32 butterflies as the real 32 x 32 GEMM
//         [Re y; Im y] = A [Re x; Im x] with v_mfma_f32_32x32x16_f16, split
//         into FP16 high and low parts (A_hi x_hi + A_hi x_lo + A_lo x_hi,
//         FP32 accumulation), which keeps FP32-level accuracy.
//
// Both apply the unitary DFT/4 in place `stages` times to 256 butterflies
// per workgroup (lines held k-major in LDS, conflict-free for both), with a
// barrier between the read and write halves of every stage.  The inter-stage
// twiddles of a real FFT are the same for both and are left out.
//
//   mfma_stage <outdir>   writes input.f32, valu.f32, mfma.f32 and
//                         mfma16.f32 (2 stages on 64 workgroups) and prints
//                         one JSON line with the timed rates (200 stages on
//                         4096 workgroups), including a one-product FP16
//                         MFMA form as the upper bound of any MFMA form
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#define CHECK(x)                                                        \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ != hipSuccess) {                                             \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));      \
      std::exit(1);                                                     \
    }                                                                   \
  } while (0)

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr int NBF = 256;  // butterflies per workgroup (one per thread)
constexpr int R = 16;     // radix

// x[k][b] complex, k-major: element (k, b) at index k * NBF + b
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
  return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// forward radix-4 DFT in place (w4 = -i)
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2,
                                     float2& a3) {
  const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y);
  const float2 t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
  const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
  const float2 d = make_float2(a1.x - a3.x, a1.y - a3.y);
  const float2 t3 = make_float2(d.y, -d.x);  // -i d
  a0 = make_float2(t0.x + t2.x, t0.y + t2.y);
  a2 = make_float2(t0.x - t2.x, t0.y - t2.y);
  a1 = make_float2(t1.x + t3.x, t1.y + t3.y);
  a3 = make_float2(t1.x - t3.x, t1.y - t3.y);
}

__global__ void __launch_bounds__(256) valu_kernel(float2* data, int stages) {
  __shared__ float2 lds[R * NBF];
  float2* g = data + (size_t)blockIdx.x * R * NBF;
  const int b = threadIdx.x;
  for (int k = 0; k < R; ++k) lds[k * NBF + b] = g[k * NBF + b];
  __syncthreads();
  // W16^m = cos(2 pi m / 16) - i sin(2 pi m / 16), m = n1 k2 for n1, k2 < 4
  // (constant indices after unrolling, so the table folds into literals)
  constexpr float kc[10] = {1.f,         0.92387953f,  0.70710678f,
                            0.38268343f, 0.f,          -0.38268343f,
                            -0.70710678f, -0.92387953f, -1.f,
                            -0.92387953f};
  constexpr float ks[10] = {0.f,         0.38268343f, 0.70710678f,
                            0.92387953f, 1.f,         0.92387953f,
                            0.70710678f, 0.38268343f, 0.f,
                            -0.38268343f};
  for (int st = 0; st < stages; ++st) {
    float2 x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = lds[k * NBF + b];
    __syncthreads();
    // X[k2 + 4 k1] = sum_n1 W4^(n1 k1) W16^(n1 k2) sum_n2 x[n1 + 4 n2] W4^(n2 k2)
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4(x[n1], x[n1 + 4], x[n1 + 8], x[n1 + 12]);
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
      for (int k2 = 1; k2 < 4; ++k2)
        x[n1 + 4 * k2] = cmul(x[n1 + 4 * k2],
                              make_float2(kc[n1 * k2], -ks[n1 * k2]));
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
      dft4(x[4 * k2], x[4 * k2 + 1], x[4 * k2 + 2], x[4 * k2 + 3]);
    // x[4 k2 + k1] now holds X[k2 + 4 k1]; scale by 1/4 (unitary)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        const float2 v = x[4 * k2 + k1];
        lds[(k2 + 4 * k1) * NBF + b] = make_float2(0.25f * v.x, 0.25f * v.y);
      }
    __syncthreads();
  }
  for (int k = 0; k < R; ++k) g[k * NBF + b] = lds[k * NBF + b];
}

// A (32 x 32, real form of DFT16 / 4): rows i < 16 give Re y_i, rows 16 + i
// give Im y_i; columns kk < 16 take Re x_kk, 16 + kk take Im x_kk.
__device__ __forceinline__ float a_entry(int i, int kk) {
  const int yi = i & 15, xk = kk & 15;
  float s, c;
  sincosf(-2.0f * 3.14159265358979f * (float)((yi * xk) & 15) / 16.0f, &s, &c);
  // W = c + i s (s = -sin(2 pi yi xk / 16)); y = W x
  float v;
  if (i < 16)
    v = kk < 16 ? c : -s;  // Re y = c xr - s xi
  else
    v = kk < 16 ? s : c;   // Im y = s xr + c xi
  return 0.25f * v;
}

// Operand layouts of v_mfma_f32_32x32x16_f16 (wave64): lane l supplies
// A[l % 32][8 (l / 32) + t] and B[8 (l / 32) + t][l % 32], t < 8, for each
// K block of 16; D[i][l % 32] is accumulator v with
// i = 8 (v / 4) + 4 (l / 32) + v % 4.
// SPLIT = true: the three-product split form; false: one FP16 product
// (x and A rounded to FP16; an upper bound on any MFMA form's speed, not
// accurate enough for the Toeplitz passes)
template <bool SPLIT>
__global__ void __launch_bounds__(256) mfma_kernel(float2* data, int stages) {
  __shared__ float2 lds[R * NBF];
  float2* g = data + (size_t)blockIdx.x * R * NBF;
  const int tid = threadIdx.x;
  for (int k = 0; k < R; ++k) lds[k * NBF + tid] = g[k * NBF + tid];
  const int lane = tid & 63, wave = tid >> 6;
  const int col = lane & 31, kb = lane >> 5;
  // A in FP16 high and low parts, for the two K blocks (Re x, Im x)
  half8 ahi[2], alo[2];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float v = a_entry(col, 16 * c + 8 * kb + t);
      const _Float16 h = (_Float16)v;
      ahi[c][t] = h;
      alo[c][t] = (_Float16)(v - (float)h);
    }
  __syncthreads();
  for (int st = 0; st < stages; ++st) {
    float16v d[2];
#pragma unroll
    for (int grp = 0; grp < 2; ++grp) {
      const int b = wave * 64 + grp * 32 + col;
      half8 bhi[2], blo[2];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float2 v = lds[(8 * kb + t) * NBF + b];
        const _Float16 hr = (_Float16)v.x, hi = (_Float16)v.y;
        bhi[0][t] = hr;
        bhi[1][t] = hi;
        if constexpr (SPLIT) {
          blo[0][t] = (_Float16)(v.x - (float)hr);
          blo[1][t] = (_Float16)(v.y - (float)hi);
        }
      }
      float16v acc = {};
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[c], bhi[c], acc, 0, 0, 0);
        if constexpr (SPLIT) {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[c], blo[c], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo[c], bhi[c], acc, 0, 0, 0);
        }
      }
      d[grp] = acc;
    }
    __syncthreads();
    float* lf = reinterpret_cast<float*>(lds);
#pragma unroll
    for (int grp = 0; grp < 2; ++grp) {
      const int b = wave * 64 + grp * 32 + col;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = 8 * (v >> 2) + 4 * kb + (v & 3);
        lf[((i & 15) * NBF + b) * 2 + (i >> 4)] = d[grp][v];
      }
    }
    __syncthreads();
  }
  for (int k = 0; k < R; ++k) g[k * NBF + tid] = lds[k * NBF + tid];
}

static void dump(const std::string& path, const std::vector<float2>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
  std::fwrite(v.data(), sizeof(float2), v.size(), f);
  std::fclose(f);
}

template <class K>
static double time_kernel(K kern, float2* d, int nblk, int stages) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), 0, 0, d, stages);  // warm
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), 0, 0, d, stages);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return ms;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: mfma_stage <outdir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  // correctness: 2 stages on 64 workgroups
  const int nchk = 64, per = R * NBF;
  std::vector<float2> h((size_t)nchk * per);
  std::mt19937 gen(12345u);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  for (auto& x : h) x = make_float2(u(gen), u(gen));
  dump(dir + "/input.f32", h);
  float2* d = nullptr;
  const int ntime = 4096, stages = 200;
  CHECK(hipMalloc(&d, sizeof(float2) * (size_t)ntime * per));
  std::vector<float2> out(h.size());
  CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(valu_kernel, dim3(nchk), dim3(256), 0, 0, d, 2);
  CHECK(hipGetLastError());
  CHECK(hipMemcpy(out.data(), d, sizeof(float2) * h.size(), hipMemcpyDeviceToHost));
  dump(dir + "/valu.f32", out);
  CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mfma_kernel<true>, dim3(nchk), dim3(256), 0, 0, d, 2);
  CHECK(hipGetLastError());
  CHECK(hipMemcpy(out.data(), d, sizeof(float2) * h.size(), hipMemcpyDeviceToHost));
  dump(dir + "/mfma.f32", out);
  CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mfma_kernel<false>, dim3(nchk), dim3(256), 0, 0, d, 2);
  CHECK(hipGetLastError());
  CHECK(hipMemcpy(out.data(), d, sizeof(float2) * h.size(), hipMemcpyDeviceToHost));
  dump(dir + "/mfma16.f32", out);
  // timing: 200 stages on 4096 workgroups (data from the same generator)
  std::vector<float2> ht((size_t)ntime * per);
  for (auto& x : ht) x = make_float2(u(gen), u(gen));
  CHECK(hipMemcpy(d, ht.data(), sizeof(float2) * ht.size(), hipMemcpyHostToDevice));
  const double tv = time_kernel(valu_kernel, d, ntime, stages);
  CHECK(hipMemcpy(d, ht.data(), sizeof(float2) * ht.size(), hipMemcpyHostToDevice));
  const double tm = time_kernel(mfma_kernel<true>, d, ntime, stages);
  CHECK(hipMemcpy(d, ht.data(), sizeof(float2) * ht.size(), hipMemcpyHostToDevice));
  const double t16 = time_kernel(mfma_kernel<false>, d, ntime, stages);
  const double pts = (double)ntime * per * stages;  // point-stages
  std::printf("{\"point_stages\": %.0f, \"valu_ms\": %.3f, \"mfma_ms\": %.3f, "
              "\"mfma_fp16_ms\": %.3f, \"valu_gpts_per_s\": %.1f, "
              "\"mfma_gpts_per_s\": %.1f, \"mfma_fp16_gpts_per_s\": %.1f, "
              "\"mfma_over_valu_time\": %.3f, "
              "\"mfma_fp16_over_valu_time\": %.3f}\n",
              pts, tv, tm, t16, pts / (tv * 1e-3) / 1e9,
              pts / (tm * 1e-3) / 1e9, pts / (t16 * 1e-3) / 1e9, tm / tv,
              t16 / tv);
  CHECK(hipFree(d));
  return 0;
}
