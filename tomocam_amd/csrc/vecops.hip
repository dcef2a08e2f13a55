// Streaming helpers: FP64-accumulated dot products, max/min, scaling.
#include <cfloat>

#include "lds_launch.h"
#include "vecops.h"

namespace tc {

__global__ void dot_partials_kernel(const float* __restrict__ a,
                                    const float* __restrict__ b, long n,
                                    double* partials) {
  __shared__ double red[16];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += (double)a[i] * (double)b[i];
  double tot = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ void sum_partials_kernel(const double* __restrict__ p, int n,
                                    double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += p[i];
  double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = tot;
}

double dot(const float* a, const float* b, long n, hipStream_t stream) {
  const int nblk = 1024;
  DevBuf parts(sizeof(double) * (nblk + 1));
  double* p = parts.as<double>();
  if (n > 0) {
    hipLaunchKernelGGL(dot_partials_kernel, dim3(nblk), dim3(256), 0, stream,
                       a, b, n, p);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, stream, p,
                       nblk, p + nblk);
    TC_HIP(hipGetLastError());
  } else {
    TC_HIP(hipMemsetAsync(p + nblk, 0, sizeof(double), stream));
  }
  double h = 0.0;
  TC_HIP(hipMemcpyAsync(&h, p + nblk, sizeof(double), hipMemcpyDeviceToHost,
                        stream));
  TC_HIP(hipStreamSynchronize(stream));
  return h;
}

// max / min seeded like DArray::max/min (src/dist_array.h:143-152)
__global__ void minmax_partials_kernel(const float* __restrict__ a, long n,
                                       float* pmax, float* pmin) {
  __shared__ float smax[256], smin[256];
  float mx = -1.0e20f, mn = 1.0e20f;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    mx = fmaxf(mx, a[i]);
    mn = fminf(mn, a[i]);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    pmin[blockIdx.x] = smin[0];
  }
}

void minmax(const float* a, long n, float* mx, float* mn,
            hipStream_t stream) {
  const int nblk = 512;
  DevBuf parts(sizeof(float) * 2 * nblk);
  float* pmax = parts.as<float>();
  float* pmin = pmax + nblk;
  hipLaunchKernelGGL(minmax_partials_kernel, dim3(nblk), dim3(256), 0, stream,
                     a, n, pmax, pmin);
  TC_HIP(hipGetLastError());
  std::vector<float> h(2 * nblk);
  TC_HIP(hipMemcpyAsync(h.data(), pmax, sizeof(float) * 2 * nblk,
                        hipMemcpyDeviceToHost, stream));
  TC_HIP(hipStreamSynchronize(stream));
  float M = -1.0e20f, m = 1.0e20f;
  for (int i = 0; i < nblk; ++i) {
    M = std::max(M, h[i]);
    m = std::min(m, h[nblk + i]);
  }
  *mx = M;
  *mn = m;
}

__global__ void scale_kernel(const float* __restrict__ a, float* out, long n,
                             float mul, float add) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    out[i] = (a[i] + add) * mul;
}

void affine(const float* a, float* out, long n, float add, float mul,
            hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_kernel, dim3(std::min<long>(4096, cdiv(n, 256))),
                     dim3(256), 0, stream, a, out, n, mul, add);
  TC_HIP(hipGetLastError());
}

__global__ void sub_kernel(const float* __restrict__ a,
                           const float* __restrict__ b, float* out, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    out[i] = a[i] - b[i];
}

void affine_sub(const float* a, const float* b, float* out, long n,
                hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(sub_kernel, dim3(std::min<long>(4096, cdiv(n, 256))),
                     dim3(256), 0, stream, a, b, out, n);
  TC_HIP(hipGetLastError());
}

__global__ void diff2_partials_kernel(const float* __restrict__ a,
                                      const float* __restrict__ b, long n,
                                      double* partials) {
  __shared__ double red[16];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const double d = (double)a[i] - (double)b[i];
    acc += d * d;
  }
  double tot = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

double diff_norm2(const float* a, const float* b, long n, hipStream_t stream) {
  const int nblk = 1024;
  DevBuf parts(sizeof(double) * (nblk + 1));
  double* p = parts.as<double>();
  hipLaunchKernelGGL(diff2_partials_kernel, dim3(nblk), dim3(256), 0, stream,
                     a, b, n, p);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, stream, p,
                     nblk, p + nblk);
  TC_HIP(hipGetLastError());
  double h = 0.0;
  TC_HIP(hipMemcpyAsync(&h, p + nblk, sizeof(double), hipMemcpyDeviceToHost,
                        stream));
  TC_HIP(hipStreamSynchronize(stream));
  return h;
}

}  // namespace tc
