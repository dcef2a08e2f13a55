// Pre/post-processing and the full mbir2 pipeline (device resident).
#pragma once

#include "mbir.h"

namespace tc {

// src/preproc.cpp:80-85: padding per side and centre offset.
void pad_amount(int nrays, float center, int* npad, int* offset);

// src/preproc.cpp:62-65,73-108: out[s][p][z] = padded[(z + off) mod N'],
// padded = in zero-padded by npad on both sides.  in [S][P][N] -> [S][P][N+2npad].
void preproc(const float* in, float* out, int S, int P, int N, int npad,
             int offset, hipStream_t stream);

// src/postproc.cpp:69-102: crop [npad, npad + n) on both axes.
void postproc(const float* in, float* out, int S, int nin, int nout,
              hipStream_t stream);

// src/padding.cpp:37-91 / src/gpu/padding.cu:119-150: symmetric zero pad by
// `padding` in total (shift padding/2) on both axes.
void pad2d(const float* in, float* out, int S, int nin, int padding,
           hipStream_t stream);

// [P][S][ncols] projections -> [S][P][ncols - drop_last] sinograms.
void projections_to_sinogram(const float* proj, float* sino, int P, int S,
                             int ncols, bool drop_last, hipStream_t stream);

struct Mbir2Options {
  int num_iters = 100;
  float sigma = 100.f;
  float p = 1.2f;
  int refresh_every = 25;
  int verbose = 0;
  // mixed precision (FP16 gradient; see Run2Options::mixed)
  bool mixed = false;
  // > 0: L from that many power-iteration steps on T (FP64 dots), starting
  // from the all-ones image, instead of the reference's max(T 1)
  int lipschitz_iters = 0;
  // time the Toeplitz slice-group sizes once per shape (Psf::tune_groups);
  // TOMOCAM_HIP_AUTOTUNE=0 turns it off process-wide
  bool autotune = true;
  // HIP-graph replay of the steady-state attempt (Run2Options::graphs)
  bool allow_graphs = true;
};

struct Mbir2Result {
  Run2Stats stats;
  double lipschitz = 0.0;
  double step = 0.0;
  double sino_norm = 0.0;
  double max_sino = 0.0;
  int npad = 0;
  int padded = 0;
};

// tomocam::mbir2 (src/mbir2.cpp:40-145): sino [S][P][N] (device) -> recon
// [S][N][N] (device); x0 [S][N][N] is the initial guess, or NULL for the
// FBP of the normalised sinogram.
Mbir2Result mbir2(const float* sino, const float* x0, float* recon, int S,
                  int P, int N, const float* h_angles, float center,
                  const Mbir2Options& o, hipStream_t stream);

// tomocam::mbir (src/mbir.cpp:38-119): the NUFFT normal-operator variant.
// Normalises by (sino - min) / (max - min), pads the given guess x0 [S][N][N]
// (required), takes the step from max Re(A^H A 1)/n^3, and runs the
// reference's run2 with gradient(y) = Re(A^H A y)/n^3 - P^T y_d + R'(y) and
// error(v) = ||P v - y_d|| (projector, not the Toeplitz expansion).  A
// sinogram with max == min returns x0 unchanged (res.padded = 0).
Mbir2Result mbir_nufft(const float* sino, const float* x0, float* recon, int S,
                       int P, int N, const float* h_angles, float center,
                       const Mbir2Options& o, hipStream_t stream);

}  // namespace tc
