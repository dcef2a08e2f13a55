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

// Extra streams for the Toeplitz passes: slice groups go round-robin over
// the caller's stream and n - 1 `aux` streams (each with its own workspace
// slab), so one group's row passes (memory-latency bound) overlap another
// group's column pass (VALU bound).
constexpr int kMaxToeplitzStreams = 3;
struct ToeplitzStreams {
  int n;  // streams in use, 2 .. kMaxToeplitzStreams
  hipStream_t aux[kMaxToeplitzStreams - 1];
  hipEvent_t start, done[kMaxToeplitzStreams - 1];
};

class Psf {
 public:
  // From a real PSF array [2n-1][2n-1] (PSF(m) at index m + n - 1) in device
  // memory.
  Psf(const float* d_psf, int n, hipStream_t stream);
  ~Psf();
  Psf(const Psf&) = delete;
  Psf& operator=(const Psf&) = delete;

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

  // The same product (mode 0) with caller-provided workspace `ws` of
  // ws_bytes (at least workspace_bytes_per_slice()): slice groups are sized
  // to fit it, one stream, and nothing is allocated.  Partials as apply().
  void apply_ws(const float* x, float* out, int nslices, double* dot_partials,
                void* ws, size_t ws_bytes, hipStream_t stream);
  size_t workspace_bytes_per_slice() const {
    return sizeof(float2) * (size_t)H_ * rp_;
  }

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

  // Mixed-precision forms (compile-time FFT lengths only; false, launching
  // nothing, otherwise).  The input is FP16 with a per-slice power-of-two
  // scale on the device: x[s] = g16[s] / gsc[s], and the spectrum is read
  // from an FP16 copy; the arithmetic is FP32.  With `wt16` the passes also
  // store their intermediates in FP16 (pass A output scaled by a power of
  // two fixed by n, pass B output with a power-of-two scale per 32 threads
  // of a column workgroup, from their maximum); without it the
  // intermediates stay FP32.  The dot partials, T x and the committed step
  // are in the input's true units.
  bool apply_half(const void* g16, const float* gsc, float* out, int nslices,
                  double* dot_partials, hipStream_t stream, bool wt16 = true);
  bool apply_commit_half(const void* g16, const float* gsc,
                         const StepCommit& c, int nslices,
                         double* dot_partials, hipStream_t stream,
                         bool wt16 = true);

  // True when this FFT length has the compile-time passes (fused commit and
  // the mixed-precision forms).
  bool has_fused() const;

  // Autotuner: times one application of the compile-time passes for each
  // candidate slice-group size (half, once and twice the 144 MB rule's) on
  // scratch data and keeps the fastest for this (n, M, precision, nslices)
  // for the rest of the process.  Returns the chosen group size, -1 if this shape
  // was tuned before, 0 if there is nothing to tune (no compile-time plan,
  // or a single slice).
  int tune_groups(int nslices, bool half, hipStream_t stream);

  // Autotuner of the FFT length (and so of the compile-time radix plan):
  // where two compile-time plans are eligible for this width, times one
  // FP32 application of each on scratch data and rebuilds the spectrum at
  // the faster length; the choice holds for this width for the rest of the
  // process (toeplitz_fft_length returns it).  Returns the length, -1 if
  // the width was tuned before (the spectrum then follows that choice), 0
  // if there is a single candidate.
  int tune_length(int nslices, hipStream_t stream);

  // Grows the workspace for `nslices` slices.
  void reserve(int nslices);

  // True when the passes for `nslices` slices (FP16 storage if `half`) run
  // on the caller's stream alone, with no second stream to order.
  bool single_stream(int nslices, bool half) const;

 private:
  // the two-stream set-up when TOMOCAM_HIP_TOEPLITZ_STREAMS (or the tuner)
  // asks for two streams for this slice count, else nullptr
  const ToeplitzStreams* streams_for(int nslices, int elem_bytes);
  ToeplitzStreams ts_{};
  bool run_half(const void* g16, const float* gsc, float* out, int nslices,
                double* dot_partials, const StepCommit* c, hipStream_t stream,
                bool wt16);
  // (re)builds the spectrum at FFT length M from the PSF array
  void build_spectrum(int M, hipStream_t stream);
  void apply_generic(const float* x, const float* b, float* out, int nslices,
                     int mode, double* dot_partials, float2* wt,
                     hipStream_t stream);
  int n_, M_, H_, rp_;
  DevBuf arr_;
  DevBuf spec_;
  DevBuf work_;
  size_t work_slices_ = 0;
  DevBuf spec16_;  // FP16 spectrum * 2^spec16_exp_, made on first use
  int spec16_exp_ = 0;
  DevBuf cinv_;    // inverse scales of pass B's FP16 output (per 32 threads)
  size_t cinv_slices_ = 0;
};

// Scales of the mixed-precision passes (see Psf::apply_half).
struct ToeplitzHalf {
  const float* gsc;  // per-slice input scale (device)
  float sinv;        // 2^-e, e the spectrum copy's exponent
  float* cinv;       // written by pass B (one per 32 threads of a column
                     // workgroup), read by pass C
  float wsc;         // pass A output scale: n * 2^16 * wsc <= 2^15
  float cbase;       // 1 / (2^e wsc), e the spectrum copy's exponent
  float pconst;      // spectral dot back to input units: 1 / (2^e wsc^2)
                     // with FP16 intermediates, else 2^-e
};

// FFT length M (>= 2n - 1) of the Toeplitz product for an n x n image.
int toeplitz_fft_length(int n);

// Number of partial sums Psf::apply writes for `nslices` slices.
int toeplitz_num_partials(const Psf& p, int nslices);

// Sums `n` doubles on the device (fixed order) into out[0].
void reduce_partials(const double* partials, long n, double* out,
                     hipStream_t stream);

}  // namespace tc
