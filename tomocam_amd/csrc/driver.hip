// Pre/post-processing kernels and the device-resident mbir2 pipeline.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <utility>

#include "driver.h"
#include "mbir.h"
#include "nufft.h"
#include "toeplitz.h"
#include "vecops.h"

namespace tc {

void pad_amount(int nrays, float center, int* npad, int* offset) {
  const int cen = (int)std::round(center);  // halves away from zero
  const int off = (nrays + 1) / 2 - cen;
  int np = static_cast<int>(0.42 * nrays) / 2;
  if (std::abs(off) > np) np = std::abs(off);
  *npad = np;
  *offset = off;
}

__global__ void preproc_kernel(const float* __restrict__ in, float* out,
                               long rows, int N, int N2, int npad, int off) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * N2) return;
  const long r = idx / N2;
  const int z = (int)(idx - r * N2);
  int zs = (z + off) % N2;
  if (zs < 0) zs += N2;
  const int zi = zs - npad;
  out[idx] = (zi >= 0 && zi < N) ? in[r * N + zi] : 0.f;
}

void preproc(const float* in, float* out, int S, int P, int N, int npad,
             int offset, hipStream_t stream) {
  const int N2 = N + 2 * npad;
  const long rows = (long)S * P;
  hipLaunchKernelGGL(preproc_kernel, dim3(cdiv(rows * N2, 256)), dim3(256), 0,
                     stream, in, out, rows, N, N2, npad, offset);
  TC_HIP(hipGetLastError());
}

__global__ void crop_kernel(const float* __restrict__ in, float* out, int S,
                            int nin, int nout, int shift) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nout * nout;
  if (idx >= per * S) return;
  const long s = idx / per;
  const int rem = (int)(idx - s * per);
  const int r = rem / nout, c = rem - (rem / nout) * nout;
  out[idx] = in[(s * nin + r + shift) * nin + c + shift];
}

void postproc(const float* in, float* out, int S, int nin, int nout,
              hipStream_t stream) {
  TC_REQUIRE(nout <= nin, "postproc: output larger than input");
  const int shift = (nin - nout) / 2;
  const long tot = (long)S * nout * nout;
  hipLaunchKernelGGL(crop_kernel, dim3(cdiv(tot, 256)), dim3(256), 0, stream,
                     in, out, S, nin, nout, shift);
  TC_HIP(hipGetLastError());
}

__global__ void pad2d_kernel(const float* __restrict__ in, float* out, int S,
                             int nin, int nout, int shift) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nout * nout;
  if (idx >= per * S) return;
  const long s = idx / per;
  const int rem = (int)(idx - s * per);
  const int r = rem / nout - shift, c = rem - (rem / nout) * nout - shift;
  out[idx] = (r >= 0 && r < nin && c >= 0 && c < nin)
                 ? in[(s * nin + r) * nin + c]
                 : 0.f;
}

void pad2d(const float* in, float* out, int S, int nin, int padding,
           hipStream_t stream) {
  const int nout = nin + std::abs(padding);
  const int shift = std::abs(padding) / 2;
  const long tot = (long)S * nout * nout;
  hipLaunchKernelGGL(pad2d_kernel, dim3(cdiv(tot, 256)), dim3(256), 0, stream,
                     in, out, S, nin, nout, shift);
  TC_HIP(hipGetLastError());
}

__global__ void fill_kernel(float* a, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

Mbir2Result mbir2(const float* sino, const float* x0, float* recon, int S,
                  int P, int N, const float* h_angles, float center,
                  const Mbir2Options& o, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && P >= 1 && N >= 2, "mbir2: bad sinogram shape");
  TC_REQUIRE(h_angles != nullptr, "mbir2: null angles");
  // the call's buffers come from (and go back to) the thread's arena; the
  // call ends with a stream synchronisation
  ArenaScope arena;
  // no autotuning (here the NUFFT spreading chunk) unless the caller allows
  AutotuneScope tune_scope(o.autotune);
  Mbir2Result res;
  const long nsino = (long)S * P * N;

  // normalise (src/mbir2.cpp:46-54)
  float mx = 0.f, mn = 0.f;
  minmax(sino, nsino, &mx, &mn, stream);
  res.max_sino = mx;
  const size_t img_bytes = sizeof(float) * (size_t)S * N * N;
  if (mx == 0.f) {
    // src/mbir2.cpp:50-53: return the guess (none: error)
    if (!x0) throw Error(TC_ERR_VALUE, "mbir2: sinogram has ZERO max value");
    TC_HIP(hipMemcpyAsync(recon, x0, img_bytes, hipMemcpyDeviceToDevice,
                          stream));
    TC_HIP(hipStreamSynchronize(stream));
    return res;
  }
  DevBuf bsino2(sizeof(float) * nsino);
  affine(sino, bsino2.as<float>(), nsino, 0.f, 1.0f / mx, stream);

  // initial guess: given, or the FBP at the unpadded width
  // (src/mbir2.cpp:56-60)
  DevBuf bx0(img_bytes);
  if (x0)
    TC_HIP(hipMemcpyAsync(bx0.as<void>(), x0, img_bytes,
                          hipMemcpyDeviceToDevice, stream));
  else
    backproject(bsino2.as<float>(), bx0.as<float>(), S, P, N, h_angles, true,
                stream);

  // pad and centre (src/mbir2.cpp:63-66)
  int npad = 0, off = 0;
  pad_amount(N, center, &npad, &off);
  const int n2 = N + 2 * npad;
  res.npad = npad;
  res.padded = n2;
  DevBuf bsinop(sizeof(float) * (size_t)S * P * n2);
  preproc(bsino2.as<float>(), bsinop.as<float>(), S, P, N, npad, off, stream);
  bsino2.release();
  DevBuf bx0p(sizeof(float) * (size_t)S * n2 * n2);
  pad2d(bx0.as<float>(), bx0p.as<float>(), S, N, n2 - N, stream);
  bx0.release();

  // b = P^T y on the padded geometry, ||y||^2 (src/mbir2.cpp:74-77)
  DevBuf bb(sizeof(float) * (size_t)S * n2 * n2);
  backproject(bsinop.as<float>(), bb.as<float>(), S, P, n2, h_angles, false,
              stream);
  res.sino_norm = dot(bsinop.as<float>(), bsinop.as<float>(),
                      (long)S * P * n2, stream);
  bsinop.release();

  // PSF and its spectrum (src/mbir2.cpp:86-97)
  const size_t n1 = 2 * (size_t)n2 - 1;
  DevBuf bpsf(sizeof(float) * n1 * n1);
  polar_psf(h_angles, P, n2, bpsf.as<float>(), stream);
  Psf psf(bpsf.as<float>(), n2, stream);
  bpsf.release();
  // FFT length (where two compile-time plans fit this width), then the
  // slice groups and stream count of the Toeplitz passes, timed once per
  // shape and precision (TOMOCAM_HIP_AUTOTUNE=0 keeps the static rules)
  {
    const char* at = std::getenv("TOMOCAM_HIP_AUTOTUNE");
    // (the mixed-precision loop keeps FP32 intermediates: same groups)
    if (o.autotune && !(at && at[0] == '0')) {
      psf.tune_length(S, stream);
      psf.tune_groups(S, false, stream);
    }
  }

  // L = max(Re(A^H A 1)/n^3) = max(T 1); the qGGMRF Hessian term never
  // reaches the host in the reference (src/mbir2.cpp:101-112,
  // src/gpu/tv_hessian.cu:53-65).  Optionally the largest eigenvalue of T
  // by power iteration instead (Rayleigh quotient after k steps).
  if (o.lipschitz_iters > 0) {
    const long nn = (long)n2 * n2;
    DevBuf bv(sizeof(float) * nn), bw(sizeof(float) * nn);
    float* v = bv.as<float>();
    float* w = bw.as<float>();
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, stream,
                       v, nn, 1.0f / (float)n2);
    TC_HIP(hipGetLastError());
    double lam = 0.0;
    for (int k = 0; k < o.lipschitz_iters; ++k) {
      psf.apply(v, nullptr, w, 1, 0, nullptr, stream);
      const double vv = dot(v, v, nn, stream), vw = dot(v, w, nn, stream);
      const double ww = dot(w, w, nn, stream);
      TC_REQUIRE(vv > 0.0 && ww > 0.0 && std::isfinite(ww),
                 "mbir2: power iteration on T degenerated");
      lam = vw / vv;
      affine(w, v, nn, 0.f, (float)(1.0 / std::sqrt(ww)), stream);
    }
    res.lipschitz = lam;
  } else {
    const long nn = (long)n2 * n2;
    DevBuf ones(sizeof(float) * nn), t1(sizeof(float) * nn);
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, stream,
                       ones.as<float>(), nn, 1.0f);
    TC_HIP(hipGetLastError());
    psf.apply(ones.as<float>(), nullptr, t1.as<float>(), 1, 0, nullptr, stream);
    float lmax = 0.f, lmin = 0.f;
    minmax(t1.as<float>(), nn, &lmax, &lmin, stream);
    res.lipschitz = lmax;
  }
  double step = 1.0 / res.lipschitz;
  if (step > 1.0) step = 1.0;
  res.step = step;

  // Optimizer::run2 (src/mbir2.cpp:139-143)
  Run2Options ro;
  ro.num_iters = o.num_iters;
  ro.step = step;
  ro.sigma = o.sigma;
  ro.p = o.p;
  ro.sino_norm = res.sino_norm;
  ro.refresh_every = o.refresh_every;
  ro.verbose = o.verbose;
  ro.mixed = o.mixed;
  ro.allow_graphs = o.allow_graphs;
  DevBuf bsol(sizeof(float) * (size_t)S * n2 * n2);
  res.stats = run2_device(psf, bx0p.as<float>(), bb.as<float>(),
                          bsol.as<float>(), S, ro, stream);

  // crop (src/mbir2.cpp:144)
  postproc(bsol.as<float>(), recon, S, n2, N, stream);
  TC_HIP(hipStreamSynchronize(stream));
  return res;
}

// Real-data front end: projections as stored in the HDF5 file, [nproj]
// [nslices][ncols], to sinograms [nslices][nproj][nout] (nout = ncols or
// ncols - 1).  The reference transposes on the host (src/hdf5/reader.h:
// 110-114) and drops the last column of even widths (DArray::dropcol,
// src/dist_array.h:258-282; test/mbir2.cpp:104-107).
__global__ void proj_to_sino_kernel(const float* __restrict__ in,
                                    float* __restrict__ out, int P, int S,
                                    int ncols, int nout, long total) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const long row = i / nout;
    const int c = (int)(i - row * nout);
    const int s = (int)(row / P);
    const int p = (int)(row - (long)s * P);
    out[i] = in[((long)p * S + s) * ncols + c];
  }
}

void projections_to_sinogram(const float* proj, float* sino, int P, int S,
                             int ncols, bool drop_last, hipStream_t stream) {
  TC_REQUIRE(P >= 1 && S >= 1 && ncols >= 2, "projections: bad shape");
  const int nout = ncols - (drop_last ? 1 : 0);
  const long total = (long)S * P * nout;
  const unsigned nblk = (unsigned)std::min<long>(cdiv(total, 256), 65536);
  hipLaunchKernelGGL(proj_to_sino_kernel, dim3(nblk), dim3(256), 0, stream,
                     proj, sino, P, S, ncols, nout, total);
  TC_HIP(hipGetLastError());
}

// out = (1 + beta) a - beta b, i.e. a + beta (a - b)
__global__ void extrapolate_kernel(const float* a, const float* b, float* out,
                                   long n, float beta) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = fmaf(beta, a[i] - b[i], a[i]);
}

// out = a - s g
__global__ void step_kernel(const float* a, const float* g, float* out,
                            long n, float s) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = fmaf(-s, g[i], a[i]);
}

Mbir2Result mbir_nufft(const float* sino, const float* x0, float* recon, int S,
                       int P, int N, const float* h_angles, float center,
                       const Mbir2Options& o, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && P >= 1 && N >= 2, "mbir: bad sinogram shape");
  TC_REQUIRE(h_angles != nullptr && x0 != nullptr,
             "mbir: angles and an initial guess are required");
  ArenaScope arena;
  Mbir2Result res;
  const long nsino = (long)S * P * N;
  const size_t img_bytes = sizeof(float) * (size_t)S * N * N;

  // normalise to [0, 1] (src/mbir.cpp:43-54)
  float mx = 0.f, mn = 0.f;
  minmax(sino, nsino, &mx, &mn, stream);
  res.max_sino = mx;
  if (mx == mn) {
    TC_HIP(hipMemcpyAsync(recon, x0, img_bytes, hipMemcpyDeviceToDevice,
                          stream));
    TC_HIP(hipStreamSynchronize(stream));
    return res;
  }
  DevBuf bsino2(sizeof(float) * nsino);
  affine(sino, bsino2.as<float>(), nsino, -mn, 1.0f / (mx - mn), stream);

  // pad and centre (src/mbir.cpp:56-60)
  int npad = 0, off = 0;
  pad_amount(N, center, &npad, &off);
  const int n2 = N + 2 * npad;
  res.npad = npad;
  res.padded = n2;
  const long ns2 = (long)S * P * n2;
  DevBuf bsinop(sizeof(float) * ns2);
  preproc(bsino2.as<float>(), bsinop.as<float>(), S, P, N, npad, off, stream);
  bsino2.release();
  const long nimg = (long)S * n2 * n2;
  const size_t vb = sizeof(float) * (size_t)nimg;
  DevBuf bsol(vb), bx(vb), by(vb), bg(vb), bc(vb), bsinoT(vb);
  pad2d(x0, bsol.as<float>(), S, N, n2 - N, stream);
  TC_HIP(hipMemcpyAsync(bx.as<void>(), bsol.as<void>(), vb,
                        hipMemcpyDeviceToDevice, stream));

  // P^T y (src/mbir.cpp:68)
  backproject(bsinop.as<float>(), bsinoT.as<float>(), S, P, n2, h_angles,
              false, stream);
  res.sino_norm = dot(bsinop.as<float>(), bsinop.as<float>(), ns2, stream);

  PolarNufft& nu = nufft_plan(h_angles, P, n2, n2);
  const float inv_n3 = (float)(1.0 / ((double)n2 * n2 * n2));
  DevBuf bcoef(sizeof(float2) * (size_t)ns2);
  DevBuf bproj(sizeof(float) * ns2);
  // Re(A^H A v)/n^3 - b (+ R'(v))  (src/gradient.cpp:35-103)
  auto gradient = [&](const float* v, float* g, int nsl, const float* b) {
    nu.type2_real(v, bcoef.as<float2>(), nsl, stream);
    nu.type1_real(bcoef.as<float2>(), g, nsl, inv_n3, stream);
    if (b) affine_sub(g, b, g, (long)nsl * n2 * n2, stream);
  };
  // ||P v - y||  (src/error.cpp:36-88)
  auto error = [&](const float* v) {
    project(v, bproj.as<float>(), S, n2, h_angles, P, stream);
    return std::sqrt(diff_norm2(bproj.as<float>(), bsinop.as<float>(), ns2,
                                stream));
  };

  // L = max(gradient(ones, 0)); the Hessian term stays on a device copy in
  // the reference (src/mbir.cpp:83-96)
  {
    const long nn = (long)n2 * n2;
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, stream,
                       by.as<float>(), nn, 1.0f);
    TC_HIP(hipGetLastError());
    gradient(by.as<float>(), bg.as<float>(), 1, nullptr);
    float lmax = 0.f, lmin = 0.f;
    minmax(bg.as<float>(), nn, &lmax, &lmin, stream);
    res.lipschitz = lmax;
  }
  double step = 1.0 / res.lipschitz;
  if (step > 1.0) step = 1.0;
  res.step = step;

  // Optimizer::run2 (src/optimize.h:96-154), restated directly: four
  // projector evaluations per attempt as in the reference
  float* sol = bsol.as<float>();
  float* x = bx.as<float>();
  float* y = by.as<float>();
  float* g = bg.as<float>();
  float* cand = bc.as<float>();
  const unsigned nb = (unsigned)cdiv(nimg, 256);
  double t = 1.0, tnew = 1.0;
  const double step0 = step;
  for (int iter = 0; iter < o.num_iters; ++iter) {
    double xerr = 0.0;
    for (;;) {
      const double beta = tnew * (1.0 / t - 1.0);
      tnew = 0.5 * (std::sqrt(t * t * t * t + 4.0 * t * t) - t * t);
      hipLaunchKernelGGL(extrapolate_kernel, dim3(nb), dim3(256), 0, stream,
                         sol, x, y, nimg, (float)beta);
      TC_HIP(hipGetLastError());
      gradient(y, g, S, bsinoT.as<float>());
      add_total_var(y, g, S, n2, n2, o.sigma, o.p, stream);
      hipLaunchKernelGGL(step_kernel, dim3(nb), dim3(256), 0, stream, y, g,
                         cand, nimg, (float)step);
      TC_HIP(hipGetLastError());
      const double fx = error(cand);
      const double fy = error(y);
      const double gy = 0.5 * step * dot(g, g, nimg, stream);
      std::swap(sol, cand);  // sol takes the candidate on every attempt
      ++res.stats.attempts;
      if (fx > fy + gy) {
        step *= 0.9;
        ++res.stats.backtracks;
        continue;
      }
      xerr = diff_norm2(sol, x, nimg, stream);
      TC_HIP(hipMemcpyAsync(x, sol, vb, hipMemcpyDeviceToDevice, stream));
      step = step0;
      t = tnew;
      ++res.stats.iterations;
      break;
    }
    const double e = error(sol);
    res.stats.last_error = e;
    res.stats.last_xerr = std::sqrt(xerr);
    if (o.verbose)
      std::fprintf(stdout, "iter: %4d, error: %5.4e, x-err: %5.4e\n", iter, e,
                   std::sqrt(xerr));
  }
  postproc(sol, recon, S, n2, N, stream);
  TC_HIP(hipStreamSynchronize(stream));
  return res;
}

}  // namespace tc
