// Host plans and the generic batched line FFT kernel.
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>

#include "fft_device.h"
#include "fft_ops.h"

namespace tc {

std::vector<int> fft_factor(int L) {
  TC_REQUIRE(L >= 1, "FFT length must be positive");
  std::vector<int> r;
  int m = L;
  int p2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++p2;
  }
  while (p2 >= 3) {
    r.push_back(8);
    p2 -= 3;
  }
  if (p2 == 2) r.push_back(4);
  if (p2 == 1) r.push_back(2);
  for (int p : {3, 5, 7}) {
    while (m % p == 0) {
      r.push_back(p);
      m /= p;
    }
  }
  TC_REQUIRE(m == 1, "FFT length " + std::to_string(L) +
                         " is not a product of 2, 3, 5, 7");
  TC_REQUIRE((int)r.size() <= FFT_MAX_STAGES, "FFT length too large");
  return r;
}

bool fft_smooth(int L) {
  if (L < 1) return false;
  for (int p : {2, 3, 5, 7})
    while (L % p == 0) L /= p;
  return L == 1;
}

int fft_next_smooth(int n, bool even) {
  int m = n < 1 ? 1 : n;
  for (;; ++m) {
    if (even && (m % 2)) continue;
    if (fft_smooth(m)) return m;
  }
}

FFTLength::FFTLength(int L) {
  ArenaPause no_arena;  // process-lifetime table, filled synchronously
  std::vector<int> r = fft_factor(L);
  plan_.L = L;
  plan_.nst = (int)r.size();
  for (int i = 0; i < plan_.nst; ++i) plan_.radix[i] = r[i];
  std::vector<float2> h(L);
  for (int m = 0; m < L; ++m) {
    double a = -2.0 * M_PI * (double)m / (double)L;
    h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  tw_.alloc(sizeof(float2) * L);
  TC_HIP(hipMemcpy(tw_.as<void>(), h.data(), sizeof(float2) * L,
                   hipMemcpyHostToDevice));
  plan_.tw = tw_.as<float2>();
}

int FFTLength::min_radix() const {
  int m = 8;
  for (int i = 0; i < plan_.nst; ++i) m = std::min(m, plan_.radix[i]);
  return plan_.nst ? m : 1;
}

int FFTLength::nb_for(int nlines, int nthr) const {
  long need = (long)nlines * (plan_.L / min_radix());
  long nb = (need + nthr - 1) / nthr;
  if (nb <= 1) return 1;
  if (nb <= 2) return 2;
  if (nb <= 4) return 4;
  if (nb <= 8) return 8;
  return 0;  // not supported
}

const FFTLength& fft_length(int L) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::unique_ptr<FFTLength>> cache;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, L);
  auto it = cache.find(key);
  if (it == cache.end())
    it = cache.emplace(key, std::make_unique<FFTLength>(L)).first;
  return *it->second;
}

// ---------------------------------------------------------------------------
// generic batched line FFT: line l of batch b starts at b*bstride + l*lstride,
// element i at + i*estride (complex elements).  `nin` leading elements are
// read (the rest are zero); the first `nout` elements are written.
// ---------------------------------------------------------------------------

template <int DIR, int NB>
__global__ void __launch_bounds__(1024)
fft_lines_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                 FFTPlan P, int lines_per_batch, int nbatch, long bstride,
                 long lstride, long estride, int nin, int nout, int lpw,
                 int ldl, int colmajor, float scale) {
  extern __shared__ float2 sm[];
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const long total_lines = (long)lines_per_batch * nbatch;
  const long l0 = (long)blockIdx.x * lpw;
  const int nl = (int)((total_lines - l0) < lpw ? (total_lines - l0) : lpw);
  const int L = P.L;

  // load (zero beyond nin)
  for (int t = tid; t < lpw * L; t += nthr) {
    int l, i;
    if (colmajor) {
      l = t % lpw;
      i = t / lpw;
    } else {
      l = t / L;
      i = t - l * L;
    }
    float2 v = make_float2(0.f, 0.f);
    if (l < nl && i < nin) {
      long gl = l0 + l;
      long b = gl / lines_per_batch;
      long ll = gl - b * lines_per_batch;
      v = in[b * bstride + ll * lstride + (long)i * estride];
    }
    sm[l * ldl + i] = v;
  }
  __syncthreads();
  fft_lds<DIR, NB>(sm, lpw, ldl, P, tid, nthr);
  for (int t = tid; t < lpw * nout; t += nthr) {
    int l, i;
    if (colmajor) {
      l = t % lpw;
      i = t / lpw;
    } else {
      l = t / nout;
      i = t - l * nout;
    }
    if (l < nl) {
      long gl = l0 + l;
      long b = gl / lines_per_batch;
      long ll = gl - b * lines_per_batch;
      out[b * bstride + ll * lstride + (long)i * estride] =
          cscale(sm[l * ldl + i], scale);
    }
  }
}

template <int DIR, int NB>
static void launch_lines_nb(const LineFFTArgs& a, const FFTLength& F, int lpw,
                            int nthr, size_t smem, unsigned nblk,
                            hipStream_t stream) {
  auto k = fft_lines_kernel<DIR, NB>;
  TC_HIP(hipFuncSetAttribute((const void*)k,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)smem));
  hipLaunchKernelGGL(k, dim3(nblk), dim3(nthr), smem, stream, a.in, a.out,
                     F.plan(), a.lines_per_batch, a.nbatch, a.bstride,
                     a.lstride, a.estride, a.nin, a.nout, lpw, F.L(),
                     a.colmajor ? 1 : 0, a.scale);
  TC_HIP(hipGetLastError());
}

void fft_lines(const LineFFTArgs& a, int dir, hipStream_t stream) {
  TC_REQUIRE(a.nin <= a.L && a.nout <= a.L, "fft_lines: nin/nout exceed L");
  {
    const char* e = std::getenv("TOMOCAM_HIP_CT_LINES");
    if (!(e && e[0] == '0') && fft_lines_ct(a, dir, stream)) return;
  }
  const FFTLength& F = fft_length(a.L);
  // lines per workgroup: enough for coalescing of column transforms, bounded
  // by 64 KiB of LDS
  int lpw = 1;
  if (a.colmajor) {
    lpw = (int)std::max<long>(1, std::min<long>(16, 65536 / (8L * a.L)));
  } else {
    lpw = (int)std::max<long>(1, std::min<long>(4, 32768 / (8L * a.L)));
  }
  int nthr = 256;
  int nb = 0;
  for (;;) {
    nb = F.nb_for(lpw, nthr);
    if (nb != 0 && nb <= 4) break;
    if (nthr < 1024) {
      nthr *= 2;
      continue;
    }
    if (lpw > 1) {
      lpw /= 2;
      continue;
    }
    nb = F.nb_for(lpw, nthr);
    TC_REQUIRE(nb != 0, "fft_lines: FFT length too large");
    break;
  }
  size_t smem = sizeof(float2) * (size_t)lpw * F.L();
  TC_REQUIRE(smem <= 160 * 1024, "fft_lines: line does not fit in LDS");
  long total = (long)a.lines_per_batch * a.nbatch;
  unsigned nblk = cdiv(total, lpw);
  if (total == 0) return;
#define TC_LAUNCH(D, B)                                             \
  if (dir == D && nb == B) {                                        \
    launch_lines_nb<D, B>(a, F, lpw, nthr, smem, nblk, stream);     \
    return;                                                         \
  }
  TC_LAUNCH(-1, 1) TC_LAUNCH(-1, 2) TC_LAUNCH(-1, 4) TC_LAUNCH(-1, 8)
  TC_LAUNCH(1, 1) TC_LAUNCH(1, 2) TC_LAUNCH(1, 4) TC_LAUNCH(1, 8)
#undef TC_LAUNCH
  throw Error(TC_ERR_INTERNAL, "fft_lines: no kernel for configuration");
}

// Column transforms restricted to the columns [0, lo) and [cols - hi, cols)
// (the others are not needed, or are zero, in the caller).
static void fft_cols_ranges(float2* data, int nbatch, int rows, int cols,
                            long row_pitch, long slice_pitch, int lo, int hi,
                            int dir, hipStream_t stream) {
  if (lo + hi >= cols) {
    lo = cols;
    hi = 0;
  }
  LineFFTArgs a;
  a.L = rows;
  a.nbatch = nbatch;
  a.bstride = slice_pitch;
  a.lstride = 1;
  a.estride = row_pitch;
  a.nin = rows;
  a.nout = rows;
  a.colmajor = true;
  if (lo > 0) {
    a.in = a.out = data;
    a.lines_per_batch = lo;
    fft_lines(a, dir, stream);
  }
  if (hi > 0) {
    a.in = a.out = data + (cols - hi);
    a.lines_per_batch = hi;
    fft_lines(a, dir, stream);
  }
}

static void fft_rows_all(float2* data, int nbatch, int rows, int cols,
                         long row_pitch, long slice_pitch, int dir,
                         hipStream_t stream) {
  LineFFTArgs a;
  a.in = data;
  a.out = data;
  a.L = cols;
  a.lines_per_batch = rows;
  a.nbatch = nbatch;
  a.bstride = slice_pitch;
  a.lstride = row_pitch;
  a.estride = 1;
  a.nin = cols;
  a.nout = cols;
  a.colmajor = false;
  fft_lines(a, dir, stream);
}

void fft2d_pruned(float2* data, int nbatch, int n, int lo, int hi, int dir,
                  bool rows_first, hipStream_t stream) {
  const long pitch = n, slice = (long)n * n;
  if (rows_first) {
    fft_rows_all(data, nbatch, n, n, pitch, slice, dir, stream);
    fft_cols_ranges(data, nbatch, n, n, pitch, slice, lo, hi, dir, stream);
  } else {
    fft_cols_ranges(data, nbatch, n, n, pitch, slice, lo, hi, dir, stream);
    fft_rows_all(data, nbatch, n, n, pitch, slice, dir, stream);
  }
}

void fft2d(float2* data, int nbatch, int rows, int cols, long row_pitch,
           long slice_pitch, int dir, hipStream_t stream) {
  LineFFTArgs a;
  // rows: contiguous lines
  a.in = data;
  a.out = data;
  a.L = cols;
  a.lines_per_batch = rows;
  a.nbatch = nbatch;
  a.bstride = slice_pitch;
  a.lstride = row_pitch;
  a.estride = 1;
  a.nin = cols;
  a.nout = cols;
  a.colmajor = false;
  fft_lines(a, dir, stream);
  // columns
  a.L = rows;
  a.lines_per_batch = cols;
  a.lstride = 1;
  a.estride = row_pitch;
  a.nin = rows;
  a.nout = rows;
  a.colmajor = true;
  fft_lines(a, dir, stream);
}

}  // namespace tc
