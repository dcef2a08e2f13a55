// FP64 forms of the hot path: the reference's T = double instantiations
// (src/mbir2.cpp:158-167, src/mbir.cpp:124, src/radon.cpp:100,
// src/iradon.cpp:107, src/gradient2.cpp:107).  Every kernel computes in
// FP64; the NUFFT uses an exponential-of-semicircle kernel of width 16, the
// width FINUFFT's rule gives for the tolerance the reference plans double
// transforms with (1e-15, src/nufft.h:219).  These kernels are written for
// exactness at the sizes the double interface is used at, not tuned like
// the FP32 path: every FFT line must fit one workgroup's LDS (lengths up to
// 8192), which bounds the padded width to 2047 (the PSF's fine grid is
// 2 (2n - 1) points).
#pragma once

#include "tc_common.h"

namespace tc {
namespace f64 {

constexpr int kMaxFFT = 8192;

// Polar-grid NUFFT on P angles x N radial samples (r_j = 2 pi (j - N/2) /
// (N - 1), src/gpu/nufft_utils.cu:37-42) with `nmodes` modes per axis.
class Nufft {
 public:
  Nufft(const double* h_angles, int P, int N, int nmodes);
  // type 1 (iflag +1): c [S][P*N] complex -> f [S][nm][nm] = sum_j c_j
  // exp(+i (k1 x_j + k2 y_j)); `real_scale` != 0 writes Re(f) * real_scale
  // to the double array `fr` instead of the complex `f`.
  void type1(const double2* c, double2* f, double* fr, double real_scale,
             int S, hipStream_t stream);
  // type 2 (iflag -1): real f [S][nm][nm] -> c [S][P*N] = sum_k f_k
  // exp(-i (k1 x_j + k2 y_j)).
  void type2_real(const double* f, double2* c, int S, hipStream_t stream);
  int nf() const { return nf_; }

 private:
  int P_, N_, nm_, nf_;
  DevBuf cs_;    // double2 [P]: cos, sin of the angles
  DevBuf corr_;  // double [nm]: (-1)^k h / psi_hat(k)
  DevBuf fine_;  // double2 [S][nf][nf]
};

// projector (src/forward.cpp:33-52) and backprojector (src/backward.cpp:
// 42-67) in FP64
void project(const double* img, double* sino, int S, int N,
             const double* h_angles, int P, hipStream_t stream);
void backproject(const double* sino, double* img, int S, int P, int N,
                 const double* h_angles, bool fbp, hipStream_t stream);
// g = Re(A^H A x)/N^3 - b (src/gradient.cpp:35-103); b may be null
void gradient_nufft(const double* x, const double* b, double* g, int S, int N,
                    const double* h_angles, int P, hipStream_t stream);

// PointSpreadFunction<double> (src/toeplitz.h:33-160): T x = (PSF * x)/n^3
// on an M x M circular grid (M smooth, >= 2n - 1).
class Psf {
 public:
  Psf(const double* h_angles, int P, int n, hipStream_t stream);
  int n() const { return n_; }
  int M() const { return M_; }
  // out[s] = T x[s] (mode 0) or T x[s] - b[s] (mode 1)
  void apply(const double* x, const double* b, double* out, int S,
             hipStream_t stream);

 private:
  int n_, M_;
  DevBuf spec_;  // double2 [M][M]: FFT of the circularly placed PSF
  DevBuf work_;  // double2 [group][M][M]
  size_t work_slices_ = 0;
};

// qGGMRF: g += sum_d w_d rho'(x_v - x_clamp(v+d)) (src/gpu/cuTV2.cu:32-76,
// rho' in FP64), and the free-boundary prior's value and gradient (the
// differentiable form's prior)
void add_total_var(const double* x, double* g, int S, int R, int C,
                   double sigma, double p, hipStream_t stream);
void add_qggmrf_grad_free(const double* x, double* g, int S, int R, int C,
                          double sigma, double p, hipStream_t stream);
double qggmrf_value_free(const double* x, int S, int R, int C, double sigma,
                         double p, hipStream_t stream);

double dot(const double* a, const double* b, long n, hipStream_t stream);
// sum (a - b)^2 (tomocam::xerror, src/xerror.cpp:65-85)
double diff_norm2(const double* a, const double* b, long n,
                  hipStream_t stream);

// tomocam::preproc<double> (src/preproc.cpp:73-108): returns the padding;
// with out == null only the padding is computed
int preproc(const double* in, double* out, int S, int P, int N, double center,
            hipStream_t stream);
// tomocam::postproc<double> (src/postproc.cpp:69-102): centred crop
void postproc(const double* in, double* out, int S, int nin, int nout,
              hipStream_t stream);
// tomocam::pad2d<double> (src/padding.cpp:37-91): symmetric zero padding
void pad2d(const double* in, double* out, int S, int nin, int padding,
           hipStream_t stream);

struct Result {
  int iterations = 0;
  int attempts = 0;
  double lipschitz = 0.0;
  double step = 0.0;
  double sino_norm = 0.0;
  double max_sino = 0.0;
  int padded = 0;
  double last_error = 0.0;
};

// tomocam::mbir2<double> (src/mbir2.cpp:40-145) with the Toeplitz operator
// above; x0 (or null for the FBP guess) and recon are [S][N][N].
Result mbir2(const double* sino, const double* x0, double* recon, int S,
             int P, int N, const double* h_angles, double center,
             int num_iters, double sigma, double p, hipStream_t stream);
// tomocam::mbir<double> (src/mbir.cpp:38-119): NUFFT normal operator,
// projector error, x0 required.
Result mbir_nufft(const double* sino, const double* x0, double* recon, int S,
                  int P, int N, const double* h_angles, double center,
                  int num_iters, double sigma, double p, hipStream_t stream);

}  // namespace f64
}  // namespace tc
