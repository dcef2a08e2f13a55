// qGGMRF regulariser and the device-resident accelerated-gradient MBIR loop.
#pragma once

#include <functional>

#include "tc_common.h"
#include "toeplitz.h"

namespace tc {

// g[v] += sum_d w_d rho'(x_v - x_clamp(v+d)) over [S][R][C] (q = 2, c = 1e-3).
// Reference: add_total_var2 (src/tv_update.cpp:41-106), qGGMRF_kernel
// (src/gpu/cuTV2.cu:32-93).
void add_total_var(const float* x, float* g, int S, int R, int C, float sigma,
                   float p, hipStream_t stream);

// sum_d w_d q / (sigma^q c) (src/gpu/potential_function.cuh:111-114).
float tv_hessian_zero(float sigma);

// Free-boundary qGGMRF prior for the differentiable form (no counterpart in
// the reference, which only has the clamped stencil above):
//   R(x) = sum over unordered in-bounds neighbour pairs {v, v+d} of
//          w_d rho(x_v - x_{v+d}),  rho(d) = |d/s|^q / (c + |d/s|^(q-p)),
// and its exact gradient dR/dx_v = sum_{d: v+d in bounds} w_d rho'(x_v -
// x_{v+d}).  In the interior this is the reference stencil; at the faces the
// pairs leaving the volume are dropped instead of clamped.
void add_qggmrf_grad_free(const float* x, float* g, int S, int R, int C,
                          float sigma, float p, hipStream_t stream);
// R(x) into the device double d_out, asynchronously on `stream`.
void qggmrf_value_free_dev(const float* x, int S, int R, int C, float sigma,
                           float p, double* d_out, hipStream_t stream);
double qggmrf_value_free(const float* x, int S, int R, int C, float sigma,
                         float p, hipStream_t stream);

// FP16 image stacks use a row-pair layout: word (s, q, c) of a
// [S][ceil(R/2)][C] array of 32-bit words holds rows 2q and 2q + 1 of
// column c as one half2 (the second half 0 past the last row), so the
// Toeplitz row pass, which transforms rows in pairs, reads both with one
// load.  Bytes of such a stack:
size_t half_pair_bytes(int S, int R, int C);

// x [S][R][C] -> FP16 x * gsc[s] in the row-pair layout, with gsc[s] the
// power of two that maps max |x[s]| into (2^11, 2^12] (1 for an all-zero
// slice).
void to_half_slices(const float* x, void* x16, float* gsc, int S, int R,
                    int C, hipStream_t stream);

struct Run2Options {
  int num_iters = 100;
  double step = 1.0;
  float sigma = 1.0f;
  float p = 1.2f;
  double sino_norm = 0.0;  // ||y||^2 of the (normalised, padded) sinogram
  // recompute T(sol), T(x) every k accepted steps: 25 keeps the drift of
  // the recurrence at FP32 rounding (against FP64 after 100 iterations at
  // 2048^2: 7.4e-7 FP32, 6.48e-6 mixed; 6.9e-7 and 6.46e-6 with 10) for 3
  // extra applications per 100 iterations instead of 9 (3.5 % faster)
  int refresh_every = 25;
  // Mixed precision: the gradient is stored in FP16 with a per-slice
  // power-of-two scale, and T g reads it (FP32 arithmetic, spectrum and
  // intermediates); the state (sol, x, T sol, T x, b) stays FP32 and the
  // dots FP64.  The step and T g use the same
  // rounded g, so the T y recurrence drifts only by FP32 rounding, as in
  // the FP32 loop; the refresh above bounds that.
  bool mixed = false;
  // Replay the steady-state attempt from HIP graphs when the environment
  // sets TOMOCAM_HIP_GRAPHS=1 (off by default: it measured slower, see
  // run2_device).  false keeps it off regardless, for callers whose other
  // threads make synchronous HIP calls while this one might be capturing.
  bool allow_graphs = true;
  int verbose = 0;         // print the reference's per-iteration line
  // optional per-iteration record: (iter, error, x-err, step, backtracks)
  std::function<void(int, double, double, double, int)> on_iter;
};

struct Run2Stats {
  int iterations = 0;
  int attempts = 0;
  int backtracks = 0;
  int toeplitz_applies = 0;
  double last_error = 0.0;
  double last_xerr = 0.0;
  int saturated = 0;  // mixed precision: slice-attempts whose FP16 g clipped
};

// Optimizer::run2 (src/optimize.h:96-154) on device-resident state.
// x0: initial guess [S][n][n] (device, read); b: P^T y [S][n][n] (device);
// out: solution [S][n][n] (device, written).
Run2Stats run2_device(Psf& psf, const float* x0, const float* b, float* out,
                      int S, const Run2Options& opt, hipStream_t stream);

// Attempts replayed from HIP graphs by run2_device in this process.
long graph_replays();

}  // namespace tc
