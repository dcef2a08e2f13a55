// Toeplitz passes on compile-time FFT plans (fft_ct.h) for the workload
// sizes; other sizes use the generic kernels in toeplitz.hip.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include <hip/hip_fp16.h>

#include "fft_ct.h"
#include "lds_launch.h"
#include "toeplitz.h"

namespace tc {

using ct::Radices;

template <int L, int E, int DIR, class Rad>
struct Xform;
template <int L, int E, int DIR, int... Rs>
struct Xform<L, E, DIR, Radices<Rs...>> {
  static constexpr int first = ct::RadixList<Rs...>::at(0);
  static constexpr int last = ct::RadixList<Rs...>::at(sizeof...(Rs) - 1);
  // ZIN: input zero from index L/2 on (every line here is zero-padded
  // from n <= L/2), so the first stage skips those terms
  template <bool ZIN = false>
  __device__ __forceinline__ static void run(float2* v, float2* line, int t,
                                             bool active,
                                             const float2* __restrict__ tw) {
    // the zero-upper DFT form needs a radix-2 first factor; other first
    // radices still skip the padding's loads and compute on zeros
    constexpr bool z = ZIN && ct::first_radix<first>() == 2;
    ct::transform<L, E, DIR, z, Rs...>(v, line, t, active, tw);
  }
};

// Zero padding: a line holds n <= L/2 data values (L >= 2n - 1, L even).
// Element r of a radix-R butterfly sits at index j + r L/R with
// j < L/R, so for r >= R/2 it lies in the padding: stage-1 inputs there are
// zero and last-stage outputs there are not kept.

// Plans: forward radices Fwd, inverse radices Inv = reversed Fwd, LPW lines
// per workgroup in the row passes and LPWC in the column pass.  ZA: pass A
// skips the zero padding of its input in the first FFT stage.
template <int L, class Rad>
struct LineElems;
template <int L, int... Rs>
struct LineElems<L, Radices<Rs...>> {
  static constexpr int value = ct::line_elems<L, Rs...>();
};

// Stride between the LDS lines of a workgroup with `lines` lines: at least
// `elems`, and congruent to ceil(32 / lines) modulo 32, so that the row
// passes' Hermitian loops (lane q on line q mod lines, element q / lines)
// spread a wave over distinct banks.
constexpr int line_stride(int elems, int lines) {
  if (lines <= 1) return elems;
  const int target = (32 + lines - 1) / lines % 32;
  return elems + (((target - elems) % 32) + 32) % 32;
}

template <int L_, int E_, int LPW_, class Fwd_, class Inv_, int LPWC_ = LPW_,
          bool ZA_ = true>
struct TPlan {
  static constexpr int L = L_, E = E_, LPW = LPW_, T = L_ / E_;
  static constexpr int NT = LPW_ * (L_ / E_);
  static constexpr int LPWC = LPWC_;
  static constexpr int NTC = LPWC_ * (L_ / E_);
  static constexpr int NTCL = (NTC + 63) / 64 * 64;  // column-pass block
  static constexpr bool ZA = ZA_;
  using Fwd = Fwd_;
  using Inv = Inv_;
  // LDS elements per line (padded exchanges) and line strides of the row
  // passes (LS) and the column pass (LSC)
  static constexpr int LE = LineElems<L_, Fwd_>::value > LineElems<L_, Inv_>::value
                                ? LineElems<L_, Fwd_>::value
                                : LineElems<L_, Inv_>::value;
  static constexpr int LS = line_stride(LE, LPW_);
  static constexpr int LSC = line_stride(LE, LPWC_);
  // FP16 column pass: one scale per 32 consecutive threads of a workgroup
  // (SG per workgroup); scales per slice for H columns
  static constexpr int SG = (LPWC_ * (L_ / E_) + 31) / 32;
  static constexpr int scales_per_slice(int H) {
    return (H + LPWC_ - 1) / LPWC_ * SG;
  }
};

// The scale group of element `row` of column k1 in the FP16 column pass's
// output: the 32-thread group of the workgroup thread that stored it
// (thread (k1 mod LPWC) T + row mod T; the inverse's last stage stores
// rows t + r L/R1, and T = L/R1 when E = R1).
template <class P>
__device__ __forceinline__ int col_scale_index(int s, int k1, int row, int H) {
  constexpr int T = P::T, LPWC = P::LPWC;
  const int kg = k1 / LPWC;
  const int tid = (k1 - kg * LPWC) * T + row % T;
  return (s * ((H + LPWC - 1) / LPWC) + kg) * P::SG + (tid >> 5);
}

using Plan750 = TPlan<750, 30, 10, Radices<10, 15, 5>, Radices<5, 15, 10>>;
using Plan1458 = TPlan<1458, 18, 3, Radices<18, 9, 9>, Radices<9, 9, 18>>;
using Plan2250 = TPlan<2250, 30, 3, Radices<10, 15, 15>, Radices<15, 15, 10>>;
#ifndef TC_P2916_E
#define TC_P2916_E 18
#endif
#ifndef TC_P2916_LPW
#define TC_P2916_LPW 3
#endif
#ifndef TC_P2916_LPWC
#define TC_P2916_LPWC 1
#endif
// Column pass: one column per workgroup measured 5 % faster than three at
// this length; at 2250 three per workgroup stay 20 % faster than one.
// Pass A without the zero-padding shortcut: with it, 32 x 1454^2 measured
// 8 % slower in pass A (at 5832 the shortcut is 3 % faster).
using Plan2916 = TPlan<2916, TC_P2916_E, TC_P2916_LPW, Radices<18, 18, 9>,
                       Radices<9, 18, 18>, TC_P2916_LPWC, false>;
using Plan4500 = TPlan<4500, 30, 2, Radices<10, 15, 30>, Radices<30, 15, 10>>;
using Plan4608 = TPlan<4608, 24, 2, Radices<24, 24, 8>, Radices<8, 24, 24>, 1>;
using Plan2304 = TPlan<2304, 24, 2, Radices<24, 24, 4>, Radices<4, 24, 24>, 2>;
using Plan768 = TPlan<768, 24, 6, Radices<24, 8, 4>, Radices<4, 8, 24>, 6>;
#ifndef TC_P5832_LPW
#define TC_P5832_LPW 1
#endif
using Plan5832 = TPlan<5832, 18, TC_P5832_LPW, Radices<18, 18, 18>,
                       Radices<18, 18, 18>, 1>;

// Wt element (slice s, column k1, row): layout [s][k1][rp], so the column
// pass reads whole columns.  (A row-group-major layout, which makes the row
// passes stream contiguously, measured slower overall: the column pass loses
// more than the row passes gain.)
__device__ __forceinline__ long wt_at(int s, int k1, int row, int H, int rp) {
  return ((long)s * H + k1) * rp + row;
}

// Line-search commit fused into passes A and C (CM = 0: none; 1: y = sol;
// 2: y = sol + beta (sol - xp)).  Pointers may alias (in-place update).
struct CommitPtrs {
  const float* sol;
  const float* xp;
  float* dst;
  const float* Tsol;
  const float* Tx;
  float* Tdst;
  float beta;
  float s;
};

// ---------------------------------------------------------------------------
// pass A: pair-packed forward row FFTs -> Wt[s][k1][row]
//   (+ dst = y - s x when committing a step: x is the gradient there)
// ---------------------------------------------------------------------------

// FP16 storage helpers of the mixed-precision passes
__device__ __forceinline__ unsigned pack_half2(float a, float b) {
  const __half2 h = __floats2half2_rn(a, b);
  return *reinterpret_cast<const unsigned*>(&h);
}
__device__ __forceinline__ float2 unpack_half2(unsigned u) {
  return __half22float2(*reinterpret_cast<const __half2*>(&u));
}

// Storage of the passes (HM bits): HM_IN16 the input x is FP16 with a
// per-slice power-of-two scale (x = x16 / gsc); HM_WT16 the intermediates
// Wt are FP16 (pass A output at a fixed scale, pass B output with a scale
// per 32 threads); HM_SPEC16 the spectrum is read from its FP16 copy.
// Arithmetic is FP32 throughout.  Forms used: 0 (FP32), HM_IN16 (the
// mixed-precision MBIR: FP16 gradient, FP32 spectrum and intermediates)
// and all three (the FP16-storage operator).
constexpr int HM_IN16 = 1, HM_WT16 = 2, HM_SPEC16 = 4;

template <class P, int CM, int HM>
__global__ void __launch_bounds__(P::NT)
ct_rows_fwd(const void* __restrict__ xv, void* __restrict__ wtv,
            const float2* __restrict__ tw, int n, int H, int rp, int nbps,
            CommitPtrs cp, ToeplitzHalf ha) {
  extern __shared__ float2 sm[];
  constexpr int L = P::L, E = P::E, T = P::T, LPW = P::LPW;
  using F = Xform<L, E, -1, typename P::Fwd>;
  constexpr int R1 = F::first, RS = F::last;
  const int line = threadIdx.x / T;
  const int t = threadIdx.x - line * T;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int s = b / nbps;
  const int r0 = (b - s * nbps) * 2 * LPW;
  const int ra = r0 + 2 * line;
  const bool hasa = ra < n, hasb = ra + 1 < n;
  // All loads first, then the commit's stores: interleaved, the compiler
  // cannot prove the stores miss later loads (dst may alias sol) and would
  // wait out each load's latency in turn.
  float2 v[E];
  float2 y[CM != 0 ? E : 1];
  const long oa = (long)s * n * n + (long)ra * n;
  // FP16 input: the row-pair layout, rows ra and ra + 1 in one word
  const unsigned* xw = reinterpret_cast<const unsigned*>(xv) +
                       ((long)s * ((n + 1) / 2) + ra / 2) * n;
  static_assert(R1 % 2 == 0, "plans need an even first radix");
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int c = t + i * T + r * (L / R1);
      const int e = i * R1 + r;
      float re = 0.f, im = 0.f;
      if ((!P::ZA || r < R1 / 2) && c < n) {
        if constexpr ((HM & HM_IN16) != 0) {
          // the raw word (second half 0 past the last row), unpacked only
          // once every load is issued: an unpack next to its load makes
          // each load wait out its latency in turn
          if (hasa) re = __uint_as_float(xw[c]);
        } else {
          if (hasa) re = reinterpret_cast<const float*>(xv)[oa + c];
          if (hasb) im = reinterpret_cast<const float*>(xv)[oa + n + c];
        }
        if constexpr (CM != 0) {
          float ya = 0.f, yb = 0.f;
          if (hasa) {
            ya = cp.sol[oa + c];
            if constexpr (CM == 2) ya = fmaf(cp.beta, ya - cp.xp[oa + c], ya);
          }
          if (hasb) {
            yb = cp.sol[oa + n + c];
            if constexpr (CM == 2)
              yb = fmaf(cp.beta, yb - cp.xp[oa + n + c], yb);
          }
          y[e] = make_float2(ya, yb);
        }
      }
      v[e] = make_float2(re, im);
    }
  if constexpr ((HM & HM_IN16) != 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = unpack_half2(__float_as_uint(v[e].x));
  }
  if constexpr (CM != 0) {
    // the step in the input's true units (FP16 input: x = g16 / gsc)
    const float se = (HM & HM_IN16) ? cp.s / ha.gsc[s] : cp.s;
#pragma unroll
    for (int i = 0; i < E / R1; ++i)
#pragma unroll
      for (int r = 0; r < (P::ZA ? R1 / 2 : R1); ++r) {
        const int c = t + i * T + r * (L / R1);
        const int e = i * R1 + r;
        if (c < n) {
          if (hasa) cp.dst[oa + c] = fmaf(-se, v[e].x, y[e].x);
          if (hasb) cp.dst[oa + n + c] = fmaf(-se, v[e].y, y[e].y);
        }
      }
  }
  float2* ln = sm + line * P::LS;
  F::template run<P::ZA>(v, ln, t, true, tw);
#pragma unroll
  for (int i = 0; i < E / RS; ++i)
#pragma unroll
    for (int r = 0; r < RS; ++r) ln[t + i * T + r * (L / RS)] = v[i * RS + r];
  __syncthreads();
  constexpr bool W16 = (HM & HM_WT16) != 0;
  const float hw = W16 ? 0.5f * ha.wsc : 0.5f;
  for (int q = threadIdx.x; q < LPW * H; q += P::NT) {
    const int k = q / LPW;
    const int l = q - k * LPW;
    const int rr = r0 + 2 * l;
    if (rr < rp) {
      const float2 zk = sm[l * P::LS + k];
      const float2 zm = cconj(sm[l * P::LS + (k == 0 ? 0 : L - k)]);
      const float2 sp = cadd(zk, zm), df = csub(zk, zm);
      if constexpr (W16) {
        // |value| <= n * 2^16 * wsc <= 2^15: no overflow
        *reinterpret_cast<uint2*>(reinterpret_cast<unsigned*>(wtv) +
                                  wt_at(s, k, rr, H, rp)) =
            make_uint2(pack_half2(hw * sp.x, hw * sp.y),
                       pack_half2(hw * df.y, -hw * df.x));
      } else {
        *reinterpret_cast<float4*>(reinterpret_cast<float2*>(wtv) +
                                   wt_at(s, k, rr, H, rp)) =
            make_float4(hw * sp.x, hw * sp.y, hw * df.y, -hw * df.x);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// pass B: column FFT, spectrum multiply (+ spectral x.Tx), inverse FFT, all
// without leaving registers between the two transforms
// ---------------------------------------------------------------------------

template <class P, int HM>
__global__ void __launch_bounds__(P::NTCL)
ct_cols(void* __restrict__ wtv, const void* __restrict__ specv,
        const float2* __restrict__ tw, int n, int H, int rp, int nbps,
        double* partials, ToeplitzHalf ha) {
  extern __shared__ float2 sm[];
  __shared__ double red[16];
  constexpr int L = P::L, E = P::E, T = P::T, LPW = P::LPWC;
  using F = Xform<L, E, -1, typename P::Fwd>;
  using I = Xform<L, E, 1, typename P::Inv>;
  constexpr int R1 = F::first, RS = F::last;
  static_assert(I::first == RS && I::last == R1, "inverse must reverse radices");
  const int line = threadIdx.x / T;
  const int t = threadIdx.x - line * T;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  // slice-minor order: the workgroups that read spectrum rows k1 for every
  // slice run back to back on one XCD, so each row is fetched once into L2
  const int nslices = gridDim.x / nbps;
  const int kg = b / nslices;
  const int s = b - kg * nslices;
  const int k1 = kg * LPW + line;
  // threads past the last line (the block is rounded up to whole waves)
  // are inactive
  const bool active = line < LPW && k1 < H;
  const int kc = active ? k1 : 0;
  // the slice's FP16 input scale (for the dot), fetched up front
  const float gsc_s = (HM & HM_IN16) ? ha.gsc[s] : 1.0f;
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int idx = t + i * T + r * (L / R1);
      float2 w = make_float2(0.f, 0.f);
      if (r < R1 / 2 && active && idx < n) {
        const long o = wt_at(s, kc, idx, H, rp);
        // FP16: the raw word, unpacked once every load is issued
        if constexpr ((HM & HM_WT16) != 0)
          w.x = __uint_as_float(reinterpret_cast<const unsigned*>(wtv)[o]);
        else
          w = reinterpret_cast<const float2*>(wtv)[o];
      }
      v[i * R1 + r] = w;
    }
  if constexpr ((HM & HM_WT16) != 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = unpack_half2(__float_as_uint(v[e].x));
  }
  // the spectrum values this thread multiplies by, loaded up front so
  // their latency overlaps the forward transform
  // (an FP16 spectrum stays packed two to a register until its use, which
  // keeps the pass within the FP32 form's register budget)
  constexpr bool S16 = (HM & HM_SPEC16) != 0;
  float hs[S16 ? 1 : E];
  __half2 hs2[S16 ? (E + 1) / 2 : 1];
  {
    const long so = (long)kc * L;
#pragma unroll
    for (int i = 0; i < E / RS; ++i)
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int e = i * RS + r;
        const long o = so + t + i * T + r * (L / RS);
        if constexpr (S16) {
          const __half h = reinterpret_cast<const __half*>(specv)[o];
          if (e & 1)
            hs2[e / 2].y = h;
          else
            hs2[e / 2].x = h;
        } else {
          hs[e] = reinterpret_cast<const float*>(specv)[o];
        }
      }
  }
  float2* ln = sm + line * P::LSC;
  F::template run<true>(v, ln, t, active, tw);
  double acc = 0.0;
  if (active) {
    const float wgt = (k1 == 0 || 2 * k1 == L) ? 1.f : 2.f;
    float facc = 0.f;
#pragma unroll
    for (int i = 0; i < E / RS; ++i)
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int e = i * RS + r;
        float h;
        if constexpr (S16)
          h = __half2float((e & 1) ? hs2[e / 2].y : hs2[e / 2].x);
        else
          h = hs[e];
        const float2 q = v[i * RS + r];
        facc += h * (q.x * q.x + q.y * q.y);
        v[i * RS + r] = cscale(q, h);
      }
    acc = (double)(wgt * facc);
  }
  // the inverse's first stage needs no exchange (same element set)
  I::run(v, ln, t, active, tw);
  if constexpr ((HM & HM_WT16) != 0) {
    // Power-of-two scale per 32 consecutive threads (a half wave): their
    // largest kept value maps below 2^14.  A shuffle reduction within the
    // half wave (the workgroup is a whole number of waves, so every lane
    // exists; inactive ones contribute 0); no LDS and no barrier.  Thread
    // t stores rows t + i T + r L/R1, i.e. the rows congruent to t mod T.
    float m = 0.f;
    if (active) {
#pragma unroll
      for (int i = 0; i < E / R1; ++i)
#pragma unroll
        for (int r = 0; r < R1 / 2; ++r) {
          const int idx = t + i * T + r * (L / R1);
          if (idx < n)
            m = fmaxf(m, fmaxf(fabsf(v[i * R1 + r].x),
                               fabsf(v[i * R1 + r].y)));
        }
    }
#pragma unroll
    for (int off = 16; off > 0; off >>= 1)
      m = fmaxf(m, __shfl_xor(m, off, 32));
    // m < 2^ex (frexp's exponent); all-zero and tiny groups keep ex >= -100
    int ex = 14;
    if (m > 0.f && m <= 3.0e38f)
      ex = max((int)((__float_as_uint(m) >> 23) & 0xff) - 126, -100);
    const float kap = __uint_as_float((unsigned)(127 + 14 - ex) << 23);
    if (active) {
#pragma unroll
      for (int i = 0; i < E / R1; ++i)
#pragma unroll
        for (int r = 0; r < R1 / 2; ++r) {
          const int idx = t + i * T + r * (L / R1);
          if (idx < n)
            reinterpret_cast<unsigned*>(wtv)[wt_at(s, kc, idx, H, rp)] =
                pack_half2(kap * v[i * R1 + r].x, kap * v[i * R1 + r].y);
        }
    }
    if ((threadIdx.x & 31) == 0 && (int)threadIdx.x < P::NTC)
      ha.cinv[((long)s * ((H + LPW - 1) / LPW) + kg) * P::SG +
              (threadIdx.x >> 5)] =
          ha.cbase * __uint_as_float((unsigned)(127 + ex - 14) << 23);
  } else if (active) {
#pragma unroll
    for (int i = 0; i < E / R1; ++i)
#pragma unroll
      for (int r = 0; r < R1 / 2; ++r) {
        const int idx = t + i * T + r * (L / R1);
        if (idx < n)
          reinterpret_cast<float2*>(wtv)[wt_at(s, kc, idx, H, rp)] =
              v[i * R1 + r];
      }
  }
  if (partials) {
    if constexpr (HM != 0) {
      // back to the input's true units (input scale, spectrum and Wt
      // scales folded into pconst)
      acc *= (double)ha.pconst / ((double)gsc_s * gsc_s);
    }
    double tot = block_sum(acc, red);
    if (threadIdx.x == 0) partials[b] = tot;
  }
}

// ---------------------------------------------------------------------------
// pass C: Hermitian completion, pair-packed inverse row FFTs, crop, epilogue
//   MODE 0: out = T x;  1: out = T x - b;
//   2 / 3: out = Ty - s T x with Ty = Tsol (2) or Tsol + beta (Tsol - Tx) (3)
// ---------------------------------------------------------------------------

template <class P, int MODE, int HM>
__global__ void __launch_bounds__(P::NT)
ct_rows_inv(const void* __restrict__ wtv, const float* bvec, float* out,
            const float2* __restrict__ tw, int n, int H, int rp, int nbps,
            CommitPtrs cp, ToeplitzHalf ha) {
  extern __shared__ float2 sm[];
  constexpr int L = P::L, E = P::E, T = P::T, LPW = P::LPW;
  using I = Xform<L, E, 1, typename P::Inv>;
  constexpr int R1 = I::first, RS = I::last;
  const int line = threadIdx.x / T;
  const int t = threadIdx.x - line * T;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int s = b / nbps;
  const int r0 = (b - s * nbps) * 2 * LPW;
  // the slice's FP16 input scale, fetched now so that its latency overlaps
  // the transform instead of delaying the epilogue
  const float gsc_s = (HM & HM_IN16) ? ha.gsc[s] : 1.0f;
  // every Wt load of the thread in flight at once, then the LDS writes
  constexpr int HC = L / 2 + 1;
  constexpr int NQ = (LPW * HC + P::NT - 1) / P::NT;
  float4 w4[NQ];
  constexpr bool W16 = (HM & HM_WT16) != 0;
  float cs[W16 ? NQ : 1], cs1[W16 ? NQ : 1];
  // FP16 Wt: raw words, unpacked after every load is issued
  uint2 u16[W16 ? NQ : 1];
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const int q = threadIdx.x + j * P::NT;
    const int k = q / LPW;
    const int rr = r0 + 2 * (q - k * LPW);
    w4[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (W16) {
      cs[j] = cs1[j] = 0.f;
      u16[j] = make_uint2(0u, 0u);
    }
    if (q < LPW * H && rr < rp) {
      const long o = wt_at(s, k, rr, H, rp);
      if constexpr (W16) {
        u16[j] = *reinterpret_cast<const uint2*>(
            reinterpret_cast<const unsigned*>(wtv) + o);
        // rows rr and rr + 1 may lie in different scale groups (odd T)
        const int i0 = col_scale_index<P>(s, k, rr, H);
        const int i1 = col_scale_index<P>(s, k, rr + 1, H);
        cs[j] = ha.cinv[i0];
        cs1[j] = i1 == i0 ? cs[j] : ha.cinv[i1];
      } else {
        w4[j] = *reinterpret_cast<const float4*>(
            reinterpret_cast<const float2*>(wtv) + o);
      }
    }
  }
  if constexpr (W16) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const float2 a = unpack_half2(u16[j].x), c = unpack_half2(u16[j].y);
      w4[j] = make_float4(a.x, a.y, c.x, c.y);
    }
  }
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const int q = threadIdx.x + j * P::NT;
    if (q < LPW * H) {
      const int k = q / LPW;
      const int l = q - k * LPW;
      float4 w = w4[j];
      if constexpr (W16) {
        w.x *= cs[j];
        w.y *= cs[j];
        w.z *= cs1[j];
        w.w *= cs1[j];
      }
      if (k == 0 || 2 * k == L) {
        w.y = 0.f;
        w.w = 0.f;
      }
      sm[l * P::LS + k] = make_float2(w.x - w.w, w.y + w.z);
      if (k > 0 && L - k >= H)
        sm[l * P::LS + L - k] = make_float2(w.x + w.w, w.z - w.y);
    }
  }
  __syncthreads();
  float2* ln = sm + line * P::LS;
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) v[i * R1 + r] = ln[t + i * T + r * (L / R1)];
  __syncthreads();
  I::run(v, ln, t, true, tw);
  const int ra = r0 + 2 * line;
  const long oa = (long)s * n * n + (long)ra * n;
  const bool hasa = ra < n, hasb = ra + 1 < n;
  // FP16 input: the transform is in units of gsc times the true values;
  // an FP16 spectrum with FP32 intermediates leaves its 2^e scale here
  // (with FP16 intermediates pass B's scales include it)
  float gi = (HM & HM_IN16) ? 1.0f / gsc_s : 1.0f;
  if constexpr ((HM & HM_SPEC16) != 0 && !W16) gi *= ha.sinv;
  // Epilogue in NB batches: each batch's operands (b, or T sol and T x) are
  // all loaded before its stores, which the compiler could not otherwise
  // move past (out may alias Tsol).  Two batches instead of one take the
  // 2916-point MODE 2 kernel from 138 to 107 VGPRs, so two workgroups fit
  // on a CU: 16 % faster at 32 x 1454^2.  (Issuing the loads before the
  // transform measured slower: more registers.)
  constexpr int NB = 2, EB = E / NB;
  static_assert(E % NB == 0, "epilogue batches must divide E");
  static_assert(RS % 2 == 0, "plans need an even last inverse radix");
  const float se = (MODE >= 2 ? cp.s : 1.0f) * gi;
#pragma unroll
  for (int h = 0; h < NB; ++h) {
    float2 e0[MODE != 0 ? EB : 1], e1[MODE == 3 ? EB : 1];
    if constexpr (MODE != 0) {
      const float* src0 = MODE == 1 ? bvec : cp.Tsol;
#pragma unroll
      for (int j = 0; j < EB; ++j) {
        const int e = h * EB + j, i = e / RS, r = e - i * RS;
        const int c = t + i * T + r * (L / RS);
        float2 p0 = make_float2(0.f, 0.f), p1 = p0;
        if (r < RS / 2 && c < n) {
          if (hasa) p0.x = src0[oa + c];
          if (hasb) p0.y = src0[oa + n + c];
          if constexpr (MODE == 3) {
            if (hasa) p1.x = cp.Tx[oa + c];
            if (hasb) p1.y = cp.Tx[oa + n + c];
          }
        }
        e0[j] = p0;
        if constexpr (MODE == 3) e1[j] = p1;
      }
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int e = h * EB + j, i = e / RS, r = e - i * RS;
      const int c = t + i * T + r * (L / RS);
      if (r < RS / 2 && c < n) {
        const float2 q = v[e];
        float2 o = q;
        if constexpr (MODE == 0) {
          if constexpr (HM != 0) {
            o.x = q.x * gi;
            o.y = q.y * gi;
          }
        } else if constexpr (MODE == 1) {
          o.x = fmaf(q.x, gi, -e0[j].x);
          o.y = fmaf(q.y, gi, -e0[j].y);
        } else {
          float2 ty = e0[j];
          if constexpr (MODE == 3) {
            ty.x = fmaf(cp.beta, ty.x - e1[j].x, ty.x);
            ty.y = fmaf(cp.beta, ty.y - e1[j].y, ty.y);
          }
          o.x = fmaf(-se, q.x, ty.x);
          o.y = fmaf(-se, q.y, ty.y);
        }
        if (hasa) out[oa + c] = o.x;
        if (hasb) out[oa + n + c] = o.y;
      }
    }
  }
}

// ---------------------------------------------------------------------------

// the 144 MB rule below
int toeplitz_rule_group(int nslices, int H, int rp, int elem_bytes) {
  const double per = (double)elem_bytes * H * rp;
  const int g = (int)std::max(1.0, std::floor(144e6 / per));
  return std::min(g, nslices);
}

// Group sizes picked by the autotuner (Psf::tune_groups), keyed by
// (device, H, rp, element bytes, slices).
static std::mutex g_group_mu;
static std::map<std::tuple<int, int, int, int, int>, int> g_group_tuned;

static std::tuple<int, int, int, int, int> group_key(int H, int rp,
                                                     int elem_bytes,
                                                     int nslices) {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  return std::make_tuple(dev, H, rp, elem_bytes, nslices);
}

// group <= 0 forgets the entry
void toeplitz_set_group(int H, int rp, int elem_bytes, int nslices,
                        int group) {
  std::lock_guard<std::mutex> lk(g_group_mu);
  const auto key = group_key(H, rp, elem_bytes, nslices);
  if (group > 0)
    g_group_tuned[key] = group;
  else
    g_group_tuned.erase(key);
}

bool toeplitz_group_tuned(int H, int rp, int elem_bytes, int nslices) {
  std::lock_guard<std::mutex> lk(g_group_mu);
  return g_group_tuned.count(group_key(H, rp, elem_bytes, nslices)) != 0;
}

// Streams for the passes (1 or 2), keyed like the group sizes.
static std::map<std::tuple<int, int, int, int, int>, int> g_streams_tuned;

void toeplitz_set_streams(int H, int rp, int elem_bytes, int nslices,
                          int streams) {
  std::lock_guard<std::mutex> lk(g_group_mu);
  const auto key = group_key(H, rp, elem_bytes, nslices);
  if (streams > 0)
    g_streams_tuned[key] = streams;
  else
    g_streams_tuned.erase(key);
}

// TOMOCAM_HIP_TOEPLITZ_STREAMS (1 .. kMaxToeplitzStreams), else the
// tuner's choice, else 1
int toeplitz_streams(int H, int rp, int elem_bytes, int nslices) {
  const char* e = std::getenv("TOMOCAM_HIP_TOEPLITZ_STREAMS");
  if (e) return std::max(1, std::min(std::atoi(e), kMaxToeplitzStreams));
  std::lock_guard<std::mutex> lk(g_group_mu);
  auto it = g_streams_tuned.find(group_key(H, rp, elem_bytes, nslices));
  return it != g_streams_tuned.end() ? it->second : 1;
}

// Slices per group of passes.  A group's Wt intermediate stays resident in
// the memory-side cache between the three passes; about 144 MB per group
// (8 slices of 1454^2 in FP32) measured 14 % faster at 32 x 1454^2 than one
// group of all slices.  Order of precedence: TOMOCAM_HIP_TOEPLITZ_SLICES
// (0: all), the autotuner's choice for this shape, the 144 MB rule.
static int toeplitz_group_slices(int nslices, int H, int rp, int elem_bytes) {
  const char* e = std::getenv("TOMOCAM_HIP_TOEPLITZ_SLICES");
  if (e) {
    const int v = std::atoi(e);
    return v > 0 ? std::min(v, nslices) : nslices;
  }
  {
    std::lock_guard<std::mutex> lk(g_group_mu);
    auto it = g_group_tuned.find(group_key(H, rp, elem_bytes, nslices));
    if (it != g_group_tuned.end()) return std::min(it->second, nslices);
  }
  return toeplitz_rule_group(nslices, H, rp, elem_bytes);
}

// Slices per group for the FP16-storage passes (sizes the pass-B scale
// buffer).
int toeplitz_half_group(int nslices, int H, int rp) {
  return toeplitz_group_slices(nslices, H, rp, 4);
}

template <class P, int HM>
static void run_ct(const void* x, const float* b, float* out, void* wt,
                   const void* spec, int n, int H, int rp, int nslices,
                   int mode, double* partials, hipStream_t stream0,
                   double img_b, double spec_b, const Psf::StepCommit* sc,
                   const ToeplitzHalf* half, const ToeplitzStreams* ts,
                   int group_cap) {
  const float2* tw = fft_length(P::L).plan().tw;
  // kernel-timer classes: the FP16-storage passes are timed apart
  constexpr bool H16 = HM != 0;  // any FP16 storage: timed apart
  constexpr bool W16 = (HM & HM_WT16) != 0;
  const int kc_fwd = H16 ? KC_TOEPLITZ_ROWS_FWD_F16 : KC_TOEPLITZ_ROWS_FWD;
  const int kc_cols = H16 ? KC_TOEPLITZ_COLS_F16 : KC_TOEPLITZ_COLS;
  const int kc_inv = H16 ? KC_TOEPLITZ_ROWS_INV_F16 : KC_TOEPLITZ_ROWS_INV;
  const size_t smem = sizeof(float2) * P::LPW * P::LS;
  const size_t smemc = sizeof(float2) * P::LPWC * P::LSC;
  const int npairs = (n + 1) / 2;
  const int nbr = (int)cdiv(npairs, P::LPW);  // row-pass blocks per slice
  const int nbc = (int)cdiv(H, P::LPWC);      // column-pass blocks per slice
  const long img = (long)n * n;
  const int cm = sc ? (sc->xp ? 2 : 1) : 0;
  const int em = cm ? cm + 1 : mode;  // epilogue mode of pass C
  int group = toeplitz_group_slices(nslices, H, rp, W16 ? 4 : 8);
  if (group_cap > 0) group = std::min(group, group_cap);
  // per-slice algorithmic bytes (FP16 storage: the input and the
  // intermediates at half width)
  const double ib = img_b / nslices, sb = spec_b / nslices;
  const double ibx = (HM & HM_IN16) ? 0.5 * ib : ib;
  const double sbw = W16 ? 0.5 * sb : sb;
  const size_t xel = (HM & HM_IN16) ? 2 : 4;
  // streams in use: group gi runs on stream gi % nst (0: the caller's) with
  // workspace slab gi % nst.  Slab k is first used by group k, so a full
  // group on it implies (k + 1) * group <= nslices: the slabs fit in the
  // nslices-slice workspace.
  const int nst = ts == nullptr
                      ? 1
                      : std::max(1, std::min(ts->n, (int)cdiv(nslices, group)));
  if (nst > 1) {
    TC_HIP(hipEventRecord(ts->start, stream0));
    for (int k = 0; k < nst - 1; ++k)
      TC_HIP(hipStreamWaitEvent(ts->aux[k], ts->start, 0));
  }
  const long wslab = (long)group * H * rp;  // workspace elements per group
  for (int s0 = 0, gi = 0; s0 < nslices; s0 += group, ++gi) {
    const int ns = std::min(group, nslices - s0);
    const long o = (long)s0 * img;
    const int k = gi % nst;
    hipStream_t stream = k ? ts->aux[k - 1] : stream0;
    void* wtg = static_cast<char*>(wt) + (long)k * wslab * (W16 ? 4 : 8);
    CommitPtrs cp{};
    if (sc)
      cp = CommitPtrs{sc->sol + o, sc->xp ? sc->xp + o : nullptr,
                      sc->dst + o, sc->Tsol + o,
                      sc->Tx ? sc->Tx + o : nullptr, sc->Tdst + o, sc->beta,
                      sc->s};
    ToeplitzHalf ha{};
    if (half) {
      ha = *half;
      ha.gsc = half->gsc + s0;
      ha.cinv = half->cinv + (long)k * group * P::scales_per_slice(H);
    }
    const void* xg = static_cast<const char*>(x) + xel * o;
    const float* bg = b ? b + o : nullptr;
    float* og = (sc ? sc->Tdst : out) + o;
    {
      ktimer_begin(kc_fwd, stream);
      auto launch = [&](auto k) {
        allow_lds((const void*)k, smem);
        hipLaunchKernelGGL(k, dim3(nbr * ns), dim3(P::NT), smem, stream, xg,
                           wtg, tw, n, H, rp, nbr, cp, ha);
      };
      if (cm == 0) launch(ct_rows_fwd<P, 0, HM>);
      else if (cm == 1) launch(ct_rows_fwd<P, 1, HM>);
      else launch(ct_rows_fwd<P, 2, HM>);
      TC_HIP(hipGetLastError());
      // g in, Wt out; committing: sol (and xp) in, dst out
      ktimer_end(kc_fwd, stream,
                 ns * (ibx + sbw + (cm ? (cm == 2 ? 3.0 : 2.0) * ib : 0.0)));
    }
    {
      auto k = ct_cols<P, HM>;
      allow_lds((const void*)k, smemc);
      ktimer_begin(kc_cols, stream);
      hipLaunchKernelGGL(k, dim3(nbc * ns), dim3(P::NTCL), smemc, stream, wtg,
                         spec, tw, n, H, rp, nbc,
                         partials ? partials + (long)s0 * nbc : nullptr, ha);
      TC_HIP(hipGetLastError());
      // Wt in and out; the spectrum is read once per group
      ktimer_end(kc_cols, stream,
                 ns * 2.0 * sbw +
                     ((HM & HM_SPEC16) ? 2.0 : 4.0) * (double)H * P::L);
    }
    {
      ktimer_begin(kc_inv, stream);
      auto launch = [&](auto k) {
        allow_lds((const void*)k, smem);
        hipLaunchKernelGGL(k, dim3(nbr * ns), dim3(P::NT), smem, stream, wtg,
                           bg, og, tw, n, H, rp, nbr, cp, ha);
      };
      if (em == 0) launch(ct_rows_inv<P, 0, HM>);
      else if (em == 1) launch(ct_rows_inv<P, 1, HM>);
      else if (em == 2) launch(ct_rows_inv<P, 2, HM>);
      else launch(ct_rows_inv<P, 3, HM>);
      TC_HIP(hipGetLastError());
      // Wt in, out written; plus b (mode 1) or Tsol (and Tx) when committing
      const double extra = em == 1 || em == 2 ? 1.0 : (em == 3 ? 2.0 : 0.0);
      ktimer_end(kc_inv, stream, ns * (sbw + ib * (1.0 + extra)));
    }
  }
  for (int k = 0; k + 1 < nst; ++k) {
    TC_HIP(hipEventRecord(ts->done[k], ts->aux[k]));
    TC_HIP(hipStreamWaitEvent(stream0, ts->done[k], 0));
  }
}

// FFT lengths with compile-time plans, ascending
extern const int kToeplitzCtSizes[] = {750, 768, 1458, 2250, 2304, 2916,
                                       4500, 4608, 5832};
// relative time per point of each plan, where two can serve one width:
// the 10 * 15 * 30 plan for 4500 (radix-5 factors, 168 VGPRs in the column
// pass) measured 1.54x the 24 * 24 * 8 plan for 4608 per point at n = 2180
// (column pass 11.7 against 7.8 ms over 20 iterations); the 10 * 15 * 15
// plan for 2250 measured 1.17x the 24 * 24 * 4 plan for 2304 at n = 1090
// (3.87 against 3.39 ms); the 10 * 15 * 5 plan for 750 1.22x the 24 * 8 * 4
// plan for 768 at n = 362 (the three passes 2.98 against 2.49 ms)
extern const float kToeplitzCtCost[] = {1.22f, 1.f,  1.f, 1.17f, 1.f,
                                        1.f,   1.54f, 1.f, 1.f};
extern const int kToeplitzCtCount = 9;

// column-pass lines per workgroup (0: no compile-time plan for M)
int toeplitz_ct_lpw(int M) {
  switch (M) {
    case 750: return Plan750::LPWC;
    case 1458: return Plan1458::LPWC;
    case 2250: return Plan2250::LPWC;
    case 2916: return Plan2916::LPWC;
    case 4500: return Plan4500::LPWC;
    case 4608: return Plan4608::LPWC;
    case 2304: return Plan2304::LPWC;
    case 768: return Plan768::LPWC;
    case 5832: return Plan5832::LPWC;
    default: return 0;
  }
}

// FP16 column-pass scales per slice (0: no compile-time plan for M)
int toeplitz_ct_scales(int M, int H) {
  switch (M) {
    case 750: return Plan750::scales_per_slice(H);
    case 1458: return Plan1458::scales_per_slice(H);
    case 2250: return Plan2250::scales_per_slice(H);
    case 2916: return Plan2916::scales_per_slice(H);
    case 4500: return Plan4500::scales_per_slice(H);
    case 4608: return Plan4608::scales_per_slice(H);
    case 2304: return Plan2304::scales_per_slice(H);
    case 768: return Plan768::scales_per_slice(H);
    case 5832: return Plan5832::scales_per_slice(H);
    default: return 0;
  }
}

template <int HM>
static bool apply_ct(int M, const void* x, const float* b, float* out,
                     void* wt, const void* spec, int n, int H, int rp,
                     int nslices, int mode, double* partials,
                     hipStream_t stream, double img_b, double spec_b,
                     const Psf::StepCommit* sc, const ToeplitzHalf* half,
                     const ToeplitzStreams* ts, int group_cap = 0) {
#define TC_RUN_CT(PL)                                                        \
  run_ct<PL, HM>(x, b, out, wt, spec, n, H, rp, nslices, mode, partials,    \
                  stream, img_b, spec_b, sc, half, ts, group_cap);          \
  return true
  switch (M) {
    case 750: TC_RUN_CT(Plan750);
    case 1458: TC_RUN_CT(Plan1458);
    case 2250: TC_RUN_CT(Plan2250);
    case 2916: TC_RUN_CT(Plan2916);
    case 4500: TC_RUN_CT(Plan4500);
    case 4608: TC_RUN_CT(Plan4608);
    case 2304: TC_RUN_CT(Plan2304);
    case 768: TC_RUN_CT(Plan768);
    case 5832: TC_RUN_CT(Plan5832);
    default: return false;
  }
#undef TC_RUN_CT
}

bool toeplitz_apply_ct(int M, const float* x, const float* b, float* out,
                       float2* wt, const float* spec, int n, int H, int rp,
                       int nslices, int mode, double* partials,
                       hipStream_t stream, double img_b, double spec_b,
                       const Psf::StepCommit* sc, const ToeplitzStreams* ts,
                       int group_cap) {
  return apply_ct<0>(M, x, b, out, wt, spec, n, H, rp, nslices, mode,
                     partials, stream, img_b, spec_b, sc, nullptr, ts,
                     group_cap);
}

bool toeplitz_apply_ct_half(int M, const void* x16, float* out, void* wt,
                            const void* spec, int n, int H, int rp,
                            int nslices, double* partials, hipStream_t stream,
                            double img_b, double spec_b,
                            const Psf::StepCommit* sc,
                            const ToeplitzHalf& half, bool wt16,
                            const ToeplitzStreams* ts) {
  if (wt16)
    return apply_ct<HM_IN16 | HM_WT16 | HM_SPEC16>(
        M, x16, nullptr, out, wt, spec, n, H, rp, nslices, 0, partials,
        stream, img_b, spec_b, sc, &half, ts);
  return apply_ct<HM_IN16>(M, x16, nullptr, out, wt, spec, n, H, rp,
                           nslices, 0, partials, stream, img_b, spec_b, sc,
                           &half, ts);
}

}  // namespace tc
