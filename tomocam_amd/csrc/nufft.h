// Polar-grid NUFFTs and the projectors built on them.
//
// Reference: NUFFT::Grid (src/nufft.h:54-192, src/gpu/nufft_utils.cu:31-72),
// nufft2d1 / nufft2d2 (src/nufft.h:230-384), project (src/forward.cpp:33-52),
// backproject (src/backward.cpp:42-67).
//
// Method: exponential-of-semicircle kernel phi(z) = exp(beta (sqrt(1-z^2)-1)),
// width w = 8, beta = 2.30 w, upsampling 2 (the parameters FINUFFT's
// published rule gives for tolerance 1.2e-7).  The fine grid of nf points per
// axis is centred (point l sits at (l - nf/2) h, h = 2 pi / nf); point
// coordinates are computed in FP64 from the analytic polar geometry (never
// stored), so no coordinate arrays live in HBM.
#pragma once

#include <memory>
#include <vector>

#include "tc_common.h"

namespace tc {

constexpr int NUFFT_W = 8;

class PolarNufft {
 public:
  // Polar grid of nproj angles x npix samples per ray; nmodes Fourier modes
  // per axis (npix for the projectors, 2 npix - 1 for the PSF).
  PolarNufft(const float* h_angles, int nproj, int npix, int nmodes);

  int nproj() const { return P_; }
  int npix() const { return N_; }
  int nmodes() const { return nm_; }
  int nf() const { return nf_; }

  // Type 1 (iflag +1): out[s][k2][k1] = scale * Re sum_j c[s][j]
  // exp(+i (k1 x_j + k2 y_j)); c is [S][P][N] complex, out [S][nm][nm].
  void type1_real(const float2* c, float* out, int S, float scale,
                  hipStream_t stream);

  // Type 2 (iflag -1): c[s][j] = sum_k f[s][k2][k1] exp(-i (k1 x_j + k2 y_j))
  // for real f [S][nm][nm]; c is [S][P][N] complex.
  void type2_real(const float* f, float2* c, int S, hipStream_t stream);

 private:
  int P_, N_, nm_, nf_;
  float beta_;
  DevBuf ax_, ay_;   // double[P]: cos/sin(theta) * nf / (N - 1)
  DevBuf corr_;      // float[nm]: (-1)^k h / psi_hat(k)
  DevBuf fine_;      // float2 [batch][nf][nf]
  int fine_batch_ = 0;
  float2* fine(int batch);
  // Spreading work list, built once per geometry: one item per (fine-grid
  // tile, chunk of at most chunk_ points) so that the tiles near the
  // centre, which every ray crosses, are split over many workgroups.
  bool work_ready_ = false;
  int nwork_ = 0, nslots_ = 0, nfix_ = 0, chunk_ = 0;
  std::vector<int> counts_;  // polar points per fine-grid tile
  DevBuf work_;  // int4: tile, first point, end point, scratch slot (-1: none)
  DevBuf fix_;   // int4: tile, first slot, slot count, 0
  void count_tiles(hipStream_t stream);
  void make_work(int chunk);
};

// Centred DFT along rays by Bluestein's algorithm:
//   V[j] = sum_z a[z] exp(sign 2 pi i (j - c)(z - c) / N),  c = floor(N/2),
// i.e. fftshift(FFT_sign(ifftshift(a))) with numpy's conventions.
class RayDft {
 public:
  RayDft(int N, int sign);
  int N() const { return N_; }
  // in: [nrays][N] real (in_real) or complex; out: complex [nrays][N], or
  // real part only (out_real).  ramp multiplies V[j] by |j - c| / N (the
  // Ram-Lak filter of src/gpu/ramlak.cu:14-31); scale multiplies the output.
  void apply(const void* in, bool in_real, void* out, bool out_real,
             long nrays, bool ramp, float scale, hipStream_t stream);

 private:
  int N_, L_;
  DevBuf chirp_;  // float2[N]: exp(sign i pi (z-c)^2 / N)
  DevBuf bhat_;   // float2[L]: FFT_L of the conjugate chirp, / L
};

// Cached NUFFT / ray-DFT plans (per device and geometry); live until exit.
// NUFFT plans hold scratch, so they are also per plan slot: each host thread
// holds its own slot while it lives, and concurrent threads never share a
// plan.
PolarNufft& nufft_plan(const float* h_angles, int nproj, int npix, int nmodes);
RayDft& ray_dft_plan(int N, int sign);

// P x (src/forward.cpp:33-52): img [S][N][N] -> sino [S][P][N].
void project(const float* img, float* sino, int S, int N, const float* h_angles,
             int P, hipStream_t stream);

// P^T y (src/backward.cpp:42-67), Ram-Lak ramp if fbp:
// sino [S][P][N] -> img [S][N][N].
void backproject(const float* sino, float* img, int S, int P, int N,
                 const float* h_angles, bool fbp, hipStream_t stream);

// Real PSF array [2n-1][2n-1] of the geometry (src/toeplitz.h:57-70).
void polar_psf(const float* h_angles, int P, int n, float* d_psf,
               hipStream_t stream);

}  // namespace tc
