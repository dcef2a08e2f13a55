// Compile-time-planned, register-resident Stockham FFT for gfx950.
//
// Plan: L = R_1 * R_2 * ... * R_S with every radix dividing E (elements per
// thread); T = L / E threads cooperate on one line.  In stage s (radix R,
// Ns = R_1 ... R_{s-1}) thread t owns butterflies j = t + i T, i < E / R:
//   inputs   x[j + r L/R]                           (r < R)
//   twiddle  w_L^(r (j mod Ns) L/(Ns R))
//   outputs  y[(j - j mod Ns) R + (j mod Ns) + r Ns]
// Each radix-R DFT (R up to 27, composite radices by Cooley-Tukey in
// registers) runs entirely in registers; consecutive stages exchange data
// through one LDS line (write outputs, barrier, read inputs, barrier).
// Stage-1 inputs t + i T + r L/R_1 and last-stage outputs t + i T + r L/R_S
// are contiguous across threads, so they are read from / written to global
// memory directly.  Running the inverse with the radices reversed makes its
// stage-1 input set equal to the forward's last-stage output set, so a
// forward-multiply-inverse sequence never leaves registers in between.
#pragma once

#include "fft_device.h"

namespace tc {
namespace ct {

// ----- constexpr trigonometry (double, for compile-time twiddles) -----

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_small(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 12; ++k) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_small(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 12; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// sin / cos of 2 pi num / den
constexpr double cx_sin2pi(long num, long den) {
  num %= den;
  if (num < 0) num += den;
  // octant reduction on the exact fraction
  const double f = (double)num / (double)den;  // in [0, 1)
  const int oct = (int)(f * 8.0);
  const double a = (f - oct / 8.0) * 2.0 * kPi;  // [0, pi/4)
  switch (oct) {
    case 0: return cx_sin_small(a);
    case 1: return cx_cos_small(kPi / 4 - a);
    case 2: return cx_cos_small(a);
    case 3: return cx_sin_small(kPi / 4 - a);
    case 4: return -cx_sin_small(a);
    case 5: return -cx_cos_small(kPi / 4 - a);
    case 6: return -cx_cos_small(a);
    default: return -cx_sin_small(kPi / 4 - a);
  }
}
constexpr double cx_cos2pi(long num, long den) {
  return cx_sin2pi(num * 4 + den, den * 4);
}

constexpr int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

// ----- radix-R DFT in registers (natural order in and out) -----

template <int R, int DIR, bool PRIME = (smallest_factor(R) == R)>
struct DftC;

// primes (and 4, 8) use the hand-written kernels
template <int DIR>
struct DftC<2, DIR, true> {
  __device__ __forceinline__ static void run(float2* a) { Dft<2, DIR>::run(a); }
};
template <int DIR>
struct DftC<3, DIR, true> {
  __device__ __forceinline__ static void run(float2* a) { Dft<3, DIR>::run(a); }
};
template <int DIR>
struct DftC<5, DIR, true> {
  __device__ __forceinline__ static void run(float2* a) { Dft<5, DIR>::run(a); }
};
template <int DIR>
struct DftC<7, DIR, true> {
  __device__ __forceinline__ static void run(float2* a) { Dft<7, DIR>::run(a); }
};
template <int DIR>
struct DftC<4, DIR, false> {
  __device__ __forceinline__ static void run(float2* a) { Dft<4, DIR>::run(a); }
};
template <int DIR>
struct DftC<8, DIR, false> {
  __device__ __forceinline__ static void run(float2* a) { Dft<8, DIR>::run(a); }
};

// composite R = R1 * R2 (R1 = smallest prime factor, or 4/8 when possible):
// n = R2 n1 + n2, k = k1 + R1 k2,
// X[k1 + R1 k2] = sum_n2 w_R2^(n2 k2) [w_R^(n2 k1) sum_n1 a[R2 n1 + n2] w_R1^(n1 k1)]
template <int R>
constexpr int first_radix() {
  return (R % 8 == 0 && R != 8) ? 8
         : (R % 4 == 0 && R != 4) ? 4
                                  : smallest_factor(R);
}

template <int R, int DIR>
struct DftC<R, DIR, false> {
  static constexpr int R1 = first_radix<R>();
  static constexpr int R2 = R / R1;
  // b * w_R^e (DIR sign), e a compile-time constant
  __device__ __forceinline__ static float2 twiddle(float2 b, long e) {
    if (e % R == 0) return b;
    const float c = (float)cx_cos2pi(e, R);
    const float s = (float)(DIR * cx_sin2pi(e, R));
    return make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
  }
  __device__ __forceinline__ static void run(float2* a) {
    float2 t[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      float2 b[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) b[n1] = a[R2 * n1 + n2];
      DftC<R1, DIR>::run(b);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        t[k1 * R2 + n2] = twiddle(b[k1], (long)n2 * k1);
    }
    finish(t, a);
  }
  // The same DFT when a[R/2 .. R-1] are zero (zero-padded input): with a
  // radix-2 first factor each radix-2 sub-DFT is (a[n2], a[n2]).
  __device__ __forceinline__ static void run_zero_upper(float2* a) {
    static_assert(R1 == 2, "zero-upper form needs a radix-2 first factor");
    float2 t[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      t[n2] = a[n2];
      t[R2 + n2] = twiddle(a[n2], n2);
    }
    finish(t, a);
  }
  // second factor: radix-R2 DFTs of the twiddled columns
  __device__ __forceinline__ static void finish(const float2* t, float2* a) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      float2 c[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) c[n2] = t[k1 * R2 + n2];
      DftC<R2, DIR>::run(c);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = c[k2];
    }
  }
};

// ----- plan -----

template <int... Rs>
struct Radices;

// true if the radices multiply to L and each divides E
template <int L, int E, int... Rs>
constexpr bool plan_valid() {
  const int a[] = {Rs...};
  long p = 1;
  for (int i = 0; i < (int)sizeof...(Rs); ++i) {
    if (E % a[i]) return false;
    p *= a[i];
  }
  return p == L && L % E == 0;
}

// radix of stage s and product of the first s radices
template <int... Rs>
struct RadixList {
  static constexpr int count = sizeof...(Rs);
  static constexpr int at(int s) {
    const int a[] = {Rs...};
    return a[s];
  }
  static constexpr int ns(int s) {
    const int a[] = {Rs...};
    int p = 1;
    for (int i = 0; i < s; ++i) p *= a[i];
    return p;
  }
};

// One stage's arithmetic on registers v[E] holding inputs v[i R + r] =
// x[j + r L/R], j = t + i T; twiddle then DFT.
// ZIN (first stage only): inputs r >= R/2 of every butterfly are zero.
template <int L, int T, int E, int R, int Ns, int DIR, bool ZIN = false>
__device__ __forceinline__ void stage_compute(float2* v, int t,
                                              const float2* __restrict__ tw) {
  static_assert(!ZIN || Ns == 1, "zero inputs only in the first stage");
  constexpr int NB = E / R;
  constexpr int step = L / (Ns * R);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (Ns > 1) {
      const int k = (t + i * T) % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const float2 w = tw[k * r * step];
        v[i * R + r] = DIR < 0 ? cmul(v[i * R + r], w) : cmulc(v[i * R + r], w);
      }
    }
    if constexpr (ZIN)
      DftC<R, DIR>::run_zero_upper(v + i * R);
    else
      DftC<R, DIR>::run(v + i * R);
  }
}

// LDS layout of one exchange (stage of radix Rp, Ns = Nsp, writing; stage
// of radix R reading).  The writer's outputs form blocks of D = Nsp Rp
// consecutive elements; every M = K D elements the layout inserts P unused
// slots: index i lives at i + (i / M) P.  With 64 dword banks, a wave's
// float2 accesses are conflict-free when its lanes' addresses are distinct
// modulo 32:
//   Nsp = 1: lane j writes j Rp + r.  An odd Rp needs nothing; an even Rp
//            gets one slot per K lanes, K = 32 / gcd(Rp, 32) (or the
//            largest K the reader allows), which staggers the lanes that
//            would share a bank;
//   Nsp > 1: lane j writes (j / Nsp)(D + P) + j mod Nsp + r Nsp, which is
//            j modulo 32 when D + P = Nsp (mod 32) (K = 1); where that P
//            would take more than a tenth of the line, modulo 16, else 0.
// The reader loads c + r L/R, and L/R is a multiple of M, so on both sides
// the padded address is a per-thread base plus a compile-time offset.
// TC_LDS_PAD=0 builds the plain Stockham layout (no padding).
#ifndef TC_LDS_PAD
#define TC_LDS_PAD 1
#endif
struct XLayout {
  int M;  // elements per padded block
  int P;  // unused slots after each block
};
constexpr int gcd_c(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}
constexpr XLayout xchg_layout(int L, int Nsp, int Rp, int R) {
  const int D = Nsp * Rp;
  if (!TC_LDS_PAD) return {D, 0};
  if (Nsp == 1) {
    if (Rp & 1) return {D, 0};
    const int want = 32 / gcd_c(Rp, 32);
    const int q = (L / R) / D;  // K must divide the reader's stride / D
    int K = 1;
    for (int k = 1; k <= want; ++k)
      if (q % k == 0) K = k;
    return {D * K, 1};
  }
  // at most ~10 % of the line in padding: else the weaker 16-slot
  // condition (32 banks), else none
  for (int mod : {32, 16}) {
    const int P = (((Nsp - D) % mod) + mod) % mod;
    if (((L - 1) / D) * P * 10 <= L) return {D, P};
  }
  return {D, 0};
}

// Writes the outputs of a stage of radix Rp (Ns = Nsp) to LDS at their
// Stockham destinations, in the layout of the exchange into radix R.
template <int L, int T, int E, int Rp, int Nsp, int R>
__device__ __forceinline__ void stage_write(const float2* v, float2* line,
                                            int t) {
  constexpr int NB = E / Rp;
  constexpr int D = Nsp * Rp;
  constexpr XLayout X = xchg_layout(L, Nsp, Rp, R);
  constexpr int K = X.M / D;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = t + i * T;
    const int k = j % Nsp;
    const int blk = j / Nsp;  // output block of D elements
    // (blk D + k + r Nsp) / M = blk / K, since k + r Nsp < D
    const int base = blk * D + k + (X.P ? (blk / K) * X.P : 0);
#pragma unroll
    for (int r = 0; r < Rp; ++r) line[base + r * Nsp] = v[i * Rp + r];
  }
}

// Reads the inputs of a radix-R stage from the exchange written by
// stage_write<L, T, E, Rp, Nsp, R>.
template <int L, int T, int E, int R, int Nsp, int Rp>
__device__ __forceinline__ void stage_read(float2* v, const float2* line,
                                           int t) {
  constexpr int NB = E / R;
  constexpr XLayout X = xchg_layout(L, Nsp, Rp, R);
  constexpr int Q = L / R;  // input stride, a multiple of M
  static_assert(Q % X.M == 0, "exchange block must divide the input stride");
  constexpr int QP = Q + (Q / X.M) * X.P;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int c = t + i * T;
    const int base = c + (X.P ? (c / X.M) * X.P : 0);
#pragma unroll
    for (int r = 0; r < R; ++r) v[i * R + r] = line[base + r * QP];
  }
}

// LDS elements one line needs: the largest padded index of any exchange of
// the plan, plus one; at least L for the natural-order uses outside the
// exchanges.
template <int L, int... Rs>
constexpr int line_elems() {
  const int a[] = {Rs...};
  int need = L, nsp = 1;
  for (int s = 1; s < (int)sizeof...(Rs); ++s) {
    const XLayout X = xchg_layout(L, nsp, a[s - 1], a[s]);
    const int e = L + ((L - 1) / X.M) * X.P;
    if (e > need) need = e;
    nsp *= a[s - 1];
  }
  return need;
}

// Runs stages s..S-1 of the plan (radices Rs in order): stage s computes on
// registers; later stages first exchange through `line`.  Every thread of
// the workgroup must call this (it contains barriers); `active` masks
// threads beyond the lines of a partial workgroup.
template <int L, int T, int E, int DIR, bool ZIN, int s, int... Rs>
struct StagesDispatch {
  using RL = RadixList<Rs...>;
  __device__ __forceinline__ static void run(float2* v, float2* line, int t,
                                             bool active,
                                             const float2* __restrict__ tw) {
    if constexpr (s < RL::count) {
      constexpr int R = RL::at(s);
      constexpr int Ns = RL::ns(s);
      if constexpr (s > 0) {
        constexpr int Rp = RL::at(s - 1);
        constexpr int Nsp = RL::ns(s - 1);
        if (active) stage_write<L, T, E, Rp, Nsp, R>(v, line, t);
        __syncthreads();
        if (active) stage_read<L, T, E, R, Nsp, Rp>(v, line, t);
        __syncthreads();
      }
      if (active) stage_compute<L, T, E, R, Ns, DIR, ZIN && s == 0>(v, t, tw);
      StagesDispatch<L, T, E, DIR, ZIN, s + 1, Rs...>::run(v, line, t, active,
                                                           tw);
    }
  }
};

// Full transform on registers.  Before: v[i R1 + r] = x[t + i T + r L/R1].
// After: v[i RS + r] = X[t + i T + r L/RS].  ZIN: the input is zero from
// index L/2 on (a zero-padded line), so v[i R1 + r] for r >= R1/2 are zero
// and are not read; needs an even R1 whose first factor is 2.
template <int L, int E, int DIR, bool ZIN, int... Rs>
__device__ __forceinline__ void transform(float2* v, float2* line, int t,
                                          bool active,
                                          const float2* __restrict__ tw) {
  static_assert(plan_valid<L, E, Rs...>(), "invalid FFT plan");
  StagesDispatch<L, L / E, E, DIR, ZIN, 0, Rs...>::run(v, line, t, active,
                                                       tw);
}

// index of register element (i, r) for a stage of radix R
template <int L, int T, int R>
__device__ __forceinline__ int elem_index(int t, int i, int r) {
  return t + i * T + r * (L / R);
}

}  // namespace ct
}  // namespace tc
