// Host-callable FFT operations built on the LDS Stockham kernels.
#pragma once

#include "tc_common.h"

namespace tc {

struct LineFFTArgs {
  const float2* in = nullptr;
  float2* out = nullptr;  // may equal `in`
  int L = 0;               // transform length (2,3,5,7-smooth)
  int lines_per_batch = 0;
  int nbatch = 1;
  long bstride = 0;   // complex elements between batches
  long lstride = 0;   // complex elements between lines of a batch
  long estride = 1;   // complex elements between elements of a line
  int nin = 0;        // leading input elements read (rest treated as zero)
  int nout = 0;       // leading output elements written
  bool colmajor = false;  // lines adjacent in memory (column transforms)
  float scale = 1.0f;
};

// dir = -1: exp(-2 pi i k n / L) (forward); +1: exp(+2 pi i k n / L); both
// unnormalised (times `scale`).  Lengths with a compile-time line plan run
// on it (fft_lines_ct.hip) unless TOMOCAM_HIP_CT_LINES=0; others on the
// runtime-plan kernel.
void fft_lines(const LineFFTArgs& a, int dir, hipStream_t stream);

// The compile-time line plans: whether L has one, the smallest planned
// length >= n (0: none), and the transform (false: no plan for a.L).
bool fft_lines_ct_has(int L);
int fft_lines_ct_length(int n);
bool fft_lines_ct(const LineFFTArgs& a, int dir, hipStream_t stream);
// Chirp-z DFT of `nrays` rays of N samples on planned length L >= 2N - 1
// (false: no plan for L): out[j] = chirp[j] IFFT_L(FFT_L(in chirp) bhat)[j]
// times `scale` (and |j - N/2| / N if `ramp`).
bool bluestein_ct(const void* in, bool in_real, void* out, bool out_real,
                  const float2* chirp, const float2* bhat, int N, int L,
                  long nrays, bool ramp, float scale, hipStream_t stream);

// In-place unnormalised 2D transform of `nbatch` complex [rows][cols] arrays.
void fft2d(float2* data, int nbatch, int rows, int cols, long row_pitch,
           long slice_pitch, int dir, hipStream_t stream);

// The same on contiguous [n][n] arrays when only the columns [0, lo) and
// [n - hi, n) matter: rows_first transforms every row, then only those
// columns (the other columns of the result are left unfinished); otherwise
// those columns are transformed first (the others must be zero on input)
// and then every row.
void fft2d_pruned(float2* data, int nbatch, int n, int lo, int hi, int dir,
                  bool rows_first, hipStream_t stream);

}  // namespace tc
