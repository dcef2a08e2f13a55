// Toeplitz (PSF-convolution) normal operator T x = sum_q x[q] PSF(a-q) / n^3.
//
// Reference: PointSpreadFunction<T> (src/toeplitz.h:33-160).  The PSF is
// placed circularly on an M x M grid (M = smallest even 2,3,5,7-smooth size
// >= 2n-1), whose spectrum is real and even.  One application is three
// passes over HBM:
//   A  pair-packed real row FFTs of the n non-zero rows -> Wt[s][k1][row]
//      (k1 < H = M/2+1, stored k1-major so pass B reads whole columns);
//   B  per k1: zero-padded column FFT, multiply by the real spectrum, inverse
//      column FFT, keep the n output rows; also sum |X|^2 * spectrum (the
//      spectral form of x . T x);
//   C  pair-packed inverse row FFTs, crop to n columns, epilogue.
#pragma once

#include "tc_common.h"

namespace tc {

class Psf {
 public:
  // From a real PSF array [2n-1][2n-1] (PSF(m) at index m + n - 1) in device
  // memory.
  Psf(const float* d_psf, int n, hipStream_t stream);

  int n() const { return n_; }
  int M() const { return M_; }
  int H() const { return H_; }
  // spectrum (scaled by 1/(n^3 M^2)), layout [H][M]
  const float* spectrum() const { return spec_.as<float>(); }
  // the real PSF array [2n-1][2n-1] the spectrum was built from
  const float* array() const { return arr_.as<float>(); }

  // out[s] = T x[s] (mode 0) or T x[s] - b[s] (mode 1), for s < nslices.
  // If `dot_partials` is non-null, per-column partial sums of x . T x are
  // written there (nslices * H doubles, deterministic order).
  void apply(const float* x, const float* b, float* out, int nslices,
             int mode, double* dot_partials, hipStream_t stream);

  // T g with a line-search step committed in the same passes (pass A reads
  // g, pass C produces T g):
  //   dst  = y - s g,      y  = sol + beta (sol - xp)
  //   Tdst = T y - s T g,  Ty = Tsol + beta (Tsol - Tx)
  // (xp == Tx == nullptr: y = sol, Ty = Tsol).  dst / Tdst may alias sol /
  // Tsol; T g itself is not stored.  Returns false, launching nothing, when
  // this FFT size has no fused passes.
  struct StepCommit {
    const float* sol;
    const float* xp;
    const float* Tsol;
    const float* Tx;
    float* dst;
    float* Tdst;
    float beta;
    float s;
  };
  bool apply_commit(const float* g, const StepCommit& c, int nslices,
                    double* dot_partials, hipStream_t stream);

  // Grows the workspace for `nslices` slices.
  void reserve(int nslices);

 private:
  int n_, M_, H_, rp_;
  DevBuf arr_;
  DevBuf spec_;
  DevBuf work_;
  size_t work_slices_ = 0;
};

// Number of partial sums Psf::apply writes for `nslices` slices.
int toeplitz_num_partials(const Psf& p, int nslices);

// Sums `n` doubles on the device (fixed order) into out[0].
void reduce_partials(const double* partials, long n, double* out,
                     hipStream_t stream);

}  // namespace tc
