// LDS-resident mixed-radix Stockham FFT building blocks for gfx950.
//
// A plan factors the line length L into radices from {2,3,4,5,7,8}.  Stage s
// with radix R (Ns = product of the earlier radices) maps butterfly j in
// [0, L/R) as
//     v[r]  = src[j + r L/R] * w^(r (j mod Ns) L/(Ns R)),   r < R
//     V     = DFT_R(v)
//     dst[(j - j mod Ns) R + (j mod Ns) + r Ns] = V[r]
// which leaves the result in natural order after the last stage (Stockham
// autosort, decimation in time).  Each thread holds its butterflies of a stage
// in registers between a read phase and a write phase, so one LDS buffer of L
// complex values per line suffices.
#pragma once

#include "tc_common.h"

namespace tc {

constexpr int FFT_MAX_STAGES = 16;

struct FFTPlan {
  int L;
  int nst;
  int radix[FFT_MAX_STAGES];
  const float2* tw;  // tw[m] = exp(-2 pi i m / L), m < L (device memory)
};

// ----- radix-R DFT kernels; DIR = -1 forward (exp(-2 pi i)), +1 inverse -----

template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {  // a * (DIR * i)
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float c = -0.5f;
    const float s = 0.86602540378443864676f;  // sin(2 pi / 3)
    float2 sum = cadd(a[1], a[2]);
    float2 dif = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x + c * sum.x, a[0].y + c * sum.y);
    float2 r = mul_i<DIR>(cscale(dif, s));
    a[0] = cadd(a[0], sum);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = mul_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
    float2 p1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    float2 p2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    float2 m1 = make_float2(a[0].x + c1 * p1.x + c2 * p2.x,
                            a[0].y + c1 * p1.y + c2 * p2.y);
    float2 m2 = make_float2(a[0].x + c2 * p1.x + c1 * p2.x,
                            a[0].y + c2 * p1.y + c1 * p2.y);
    float2 r1 = mul_i<DIR>(make_float2(s1 * d1.x + s2 * d2.x,
                                       s1 * d1.y + s2 * d2.y));
    float2 r2 = mul_i<DIR>(make_float2(s2 * d1.x - s1 * d2.x,
                                       s2 * d1.y - s1 * d2.y));
    a[0] = cadd(a[0], cadd(p1, p2));
    a[1] = cadd(m1, r1);
    a[4] = csub(m1, r1);
    a[2] = cadd(m2, r2);
    a[3] = csub(m2, r2);
  }
};

template <int DIR>
struct Dft<7, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    // cos / sin of 2 pi k / 7, k = 1..3
    const float c[3] = {0.62348980185873353053f, -0.22252093395631440429f,
                        -0.90096886790241912624f};
    const float s[3] = {0.78183148246802980871f, 0.97492791218182360702f,
                        0.43388373911755812048f};
    float2 p[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      p[k] = cadd(a[k + 1], a[6 - k]);
      d[k] = csub(a[k + 1], a[6 - k]);
    }
    float2 out[7];
    out[0] = cadd(a[0], cadd(p[0], cadd(p[1], p[2])));
#pragma unroll
    for (int q = 1; q <= 3; ++q) {
      float2 m = a[0];
      float2 r = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 1; k <= 3; ++k) {
        int e = (k * q) % 7;  // angle index of w^(k q)
        float ce = (e <= 3) ? c[e - 1] : c[7 - e - 1];
        float se = (e <= 3) ? s[e - 1] : -s[7 - e - 1];
        m.x += ce * p[k - 1].x;
        m.y += ce * p[k - 1].y;
        r.x += se * d[k - 1].x;
        r.y += se * d[k - 1].y;
      }
      float2 ri = mul_i<DIR>(r);
      out[q] = cadd(m, ri);
      out[7 - q] = csub(m, ri);
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) a[q] = out[q];
  }
};

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* a) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, DIR>::run(e);
    Dft<4, DIR>::run(o);
    // o[q] *= w8^q, w8 = exp(DIR i pi / 4)
    float2 o1 = o[1];
    o[1] = make_float2(h * (o1.x - DIR * o1.y), h * (o1.y + DIR * o1.x));
    o[2] = mul_i<DIR>(o[2]);
    float2 o3 = o[3];
    o[3] = make_float2(h * (-o3.x - DIR * o3.y), h * (-o3.y + DIR * o3.x));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = cadd(e[q], o[q]);
      a[q + 4] = csub(e[q], o[q]);
    }
  }
};

// ----- one Stockham stage over `nlines` lines held in LDS -----

template <int R, int DIR, int NB>
__device__ __forceinline__ void fft_stage(float2* buf, int nlines, int ldl,
                                          int L, int Ns, const float2* tw,
                                          int tid, int nthr) {
  const int nb = L / R;
  const int total = nlines * nb;
  const int twstep = L / (Ns * R);
  float2 v[NB][R];
  int dsti[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    int t = tid + i * nthr;
    dsti[i] = -1;
    if (t < total) {
      int line = (nlines == 1) ? 0 : t / nb;
      int j = t - line * nb;
      int k = j % Ns;
      const float2* src = buf + line * ldl + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = src[r * nb];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * twstep];
          v[i][r] = (DIR < 0) ? cmul(v[i][r], w) : cmulc(v[i][r], w);
        }
      }
      Dft<R, DIR>::run(v[i]);
      dsti[i] = line * ldl + (j - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (dsti[i] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[dsti[i] + r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

// Full transform of `nlines` lines (line l at buf + l*ldl).  Must be called
// by every thread of the block with the data already in LDS and a barrier
// behind it; returns after a barrier.  NB must satisfy
// nlines * L / min_radix <= NB * nthr.
template <int DIR, int NB>
__device__ void fft_lds(float2* buf, int nlines, int ldl, const FFTPlan& P,
                        int tid, int nthr) {
  int Ns = 1;
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.radix[s];
    switch (R) {
      case 2: fft_stage<2, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
      case 3: fft_stage<3, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
      case 4: fft_stage<4, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
      case 5: fft_stage<5, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
      case 7: fft_stage<7, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
      default: fft_stage<8, DIR, NB>(buf, nlines, ldl, P.L, Ns, P.tw, tid, nthr); break;
    }
    Ns *= R;
  }
}

// ----- host side -----

// Factors L into radices {8,4,2,3,5,7}; throws if L has another prime factor.
std::vector<int> fft_factor(int L);
// True if L is a product of 2, 3, 5, 7.
bool fft_smooth(int L);
// Smallest 2,3,5,7-smooth integer >= n (even if `even`).
int fft_next_smooth(int n, bool even);

// Owns the twiddle table of one length on the current device.
class FFTLength {
 public:
  FFTLength() = default;
  explicit FFTLength(int L);
  const FFTPlan& plan() const { return plan_; }
  int L() const { return plan_.L; }
  int min_radix() const;
  // Number of butterflies per thread (1, 2, 4 or 8) needed to run `nlines`
  // lines with `nthr` threads.
  int nb_for(int nlines, int nthr) const;

 private:
  FFTPlan plan_{};
  DevBuf tw_;
};

// Cache of FFTLength objects per (device, L); entries live until exit.
const FFTLength& fft_length(int L);

}  // namespace tc
