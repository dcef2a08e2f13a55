// Batched line FFTs on compile-time plans (fft_ct.h) for the lengths the
// NUFFT fine grids and the PSF spectra use.  One workgroup transforms LPW
// whole lines: every element is loaded into registers, the stages exchange
// through one padded LDS line per line, and the outputs are stored from
// registers, so a line is read and written once.  Lengths without a plan
// fall back to the runtime-plan kernel in fft.hip.
#include <algorithm>

#include "fft_ct.h"
#include "fft_ops.h"
#include "lds_launch.h"

namespace tc {

namespace {

using ct::Radices;

// Line stride (elements) for `lines` lines of `elems` padded elements:
// congruent to ceil(32 / lines) modulo 32, so that the lines of a
// workgroup start on different LDS banks.
constexpr int lines_stride(int elems, int lines) {
  if (lines <= 1) return elems;
  const int target = (32 + lines - 1) / lines % 32;
  return elems + (((target - elems) % 32) + 32) % 32;
}

template <class Rad>
struct RadixOps;
template <int... Rs>
struct RadixOps<Radices<Rs...>> {
  static constexpr int first = ct::RadixList<Rs...>::at(0);
  static constexpr int last = ct::RadixList<Rs...>::at(sizeof...(Rs) - 1);
  template <int L>
  static constexpr int elems() { return ct::line_elems<L, Rs...>(); }
  template <int L, int E, int DIR>
  __device__ __forceinline__ static void run(float2* v, float2* line, int t,
                                             bool active,
                                             const float2* __restrict__ tw) {
    ct::transform<L, E, DIR, false, Rs...>(v, line, t, active, tw);
  }
};

// Fwd: the radices of the line transforms (either direction); Inv: the
// same radices reversed, for a forward-multiply-inverse sequence held in
// registers (the inverse's first-stage inputs are the forward's outputs).
template <int L_, int E_, int LPW_, class Fwd, class Inv>
struct LinePlan {
  static constexpr int L = L_, E = E_, LPW = LPW_, T = L_ / E_;
  static constexpr int NT = LPW_ * T;
  static constexpr int LE = RadixOps<Fwd>::template elems<L_>() >
                                    RadixOps<Inv>::template elems<L_>()
                                ? RadixOps<Fwd>::template elems<L_>()
                                : RadixOps<Inv>::template elems<L_>();
  static constexpr int LS = lines_stride(LE, LPW_);
  static constexpr int R1 = RadixOps<Fwd>::first;
  static constexpr int RS = RadixOps<Fwd>::last;
  static_assert(RadixOps<Inv>::first == RS && RadixOps<Inv>::last == R1,
                "Inv must be Fwd reversed");
  template <int DIR>
  __device__ __forceinline__ static void run(float2* v, float2* line, int t,
                                             bool active,
                                             const float2* __restrict__ tw) {
    RadixOps<Fwd>::template run<L_, E_, DIR>(v, line, t, active, tw);
  }
  __device__ __forceinline__ static void run_inv(float2* v, float2* line,
                                                 int t, bool active,
                                                 const float2* __restrict__ tw) {
    RadixOps<Inv>::template run<L_, E_, 1>(v, line, t, active, tw);
  }
};

// Line l of batch b starts at b * bstride + l * lstride; element i at
// + i * estride.  COL: the workgroup's lines are adjacent columns, so
// consecutive threads take consecutive lines (coalesced loads of a row);
// otherwise consecutive threads take consecutive elements of one line.
template <class P, int DIR, bool COL>
__global__ void __launch_bounds__(P::NT)
ct_lines_kernel(const float2* in, float2* out, const float2* __restrict__ tw,
                int lines_per_batch, long nlines, long bstride, long lstride,
                long estride, int nin, int nout, float scale) {
  extern __shared__ float2 sm[];
  constexpr int L = P::L, E = P::E, T = P::T, LPW = P::LPW;
  constexpr int R1 = P::R1, RS = P::RS;
  const int line = COL ? (int)(threadIdx.x % LPW) : (int)(threadIdx.x / T);
  const int t = COL ? (int)(threadIdx.x / LPW) : (int)(threadIdx.x % T);
  const long gl = (long)xcd_remap(blockIdx.x, gridDim.x) * LPW + line;
  const bool active = gl < nlines;
  const long b = active ? gl / lines_per_batch : 0;
  const long base = active ? b * bstride + (gl - b * lines_per_batch) * lstride
                           : 0;
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int idx = t + i * T + r * (L / R1);
      v[i * R1 + r] = (active && idx < nin) ? in[base + idx * estride]
                                            : make_float2(0.f, 0.f);
    }
  // every load of the workgroup precedes the transform's first barrier,
  // and the workgroup's lines are its own: in == out is safe
  P::template run<DIR>(v, sm + line * P::LS, t, active, tw);
  if (!active) return;
#pragma unroll
  for (int i = 0; i < E / RS; ++i)
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int idx = t + i * T + r * (L / RS);
      if (idx < nout) {
        const float2 q = v[i * RS + r];
        out[base + idx * estride] = make_float2(q.x * scale, q.y * scale);
      }
    }
}

template <class P, int DIR, bool COL>
void launch_ct_lines(const LineFFTArgs& a, hipStream_t stream) {
  const long nlines = (long)a.lines_per_batch * a.nbatch;
  if (nlines == 0) return;
  const size_t smem = sizeof(float2) * (size_t)P::LPW * P::LS;
  auto k = ct_lines_kernel<P, DIR, COL>;
  allow_lds((const void*)k, smem);
  const unsigned nblk = (unsigned)cdiv(nlines, P::LPW);
  hipLaunchKernelGGL(k, dim3(nblk), dim3(P::NT), smem, stream, a.in, a.out,
                     fft_length(P::L).plan().tw, a.lines_per_batch, nlines,
                     a.bstride, a.lstride, a.estride, a.nin, a.nout, a.scale);
  TC_HIP(hipGetLastError());
}

template <class P>
bool run_plan(const LineFFTArgs& a, int dir, hipStream_t stream) {
  if (a.colmajor) {
    if (dir < 0) launch_ct_lines<P, -1, true>(a, stream);
    else launch_ct_lines<P, 1, true>(a, stream);
  } else {
    if (dir < 0) launch_ct_lines<P, -1, false>(a, stream);
    else launch_ct_lines<P, 1, false>(a, stream);
  }
  return true;
}

// Plans: every radix divides E; LPW lines per workgroup keep it near 256
// threads where the LDS allows.
using L512 = LinePlan<512, 8, 4, Radices<8, 8, 8>, Radices<8, 8, 8>>;
using L768 = LinePlan<768, 24, 8, Radices<24, 8, 4>, Radices<4, 8, 24>>;
using L1024 = LinePlan<1024, 16, 4, Radices<16, 16, 4>, Radices<4, 16, 16>>;
using L1458 = LinePlan<1458, 18, 3, Radices<18, 9, 9>, Radices<9, 9, 18>>;
using L1536 = LinePlan<1536, 24, 4, Radices<24, 8, 8>, Radices<8, 8, 24>>;
using L2048 = LinePlan<2048, 16, 2, Radices<16, 16, 8>, Radices<8, 16, 16>>;
using L2304 = LinePlan<2304, 24, 2, Radices<24, 24, 4>, Radices<4, 24, 24>>;
using L2916 = LinePlan<2916, 18, 2, Radices<18, 18, 9>, Radices<9, 18, 18>>;
using L3072 =
    LinePlan<3072, 24, 2, Radices<24, 8, 4, 4>, Radices<4, 4, 8, 24>>;
using L4374 =
    LinePlan<4374, 18, 1, Radices<18, 9, 9, 3>, Radices<3, 9, 9, 18>>;
using L4608 = LinePlan<4608, 24, 1, Radices<24, 24, 8>, Radices<8, 24, 24>>;
using L5832 = LinePlan<5832, 18, 1, Radices<18, 18, 18>, Radices<18, 18, 18>>;
using L6144 =
    LinePlan<6144, 24, 1, Radices<24, 8, 8, 4>, Radices<4, 8, 8, 24>>;
using L8748 =
    LinePlan<8748, 18, 1, Radices<18, 18, 9, 3>, Radices<3, 9, 18, 18>>;
using L11664 =
    LinePlan<11664, 36, 1, Radices<36, 18, 18>, Radices<18, 18, 36>>;

constexpr int kLineLengths[] = {512,  768,  1024, 1458, 1536,
                                2048, 2304, 2916, 3072, 4374,
                                4608, 5832, 6144, 8748, 11664};

// Chirp-z (Bluestein) DFT of rays of N samples on a planned length
// L >= 2N - 1: out[j] = chirp[j] IFFT_L(FFT_L(in chirp) bhat)[j], j < N
// (bhat = FFT_L of the conjugate chirp / L).  The forward transform, the
// multiply and the inverse (radices reversed, so its inputs are the
// forward's outputs) stay in registers; one ray per line, LPW rays per
// workgroup.  flags: 1 real input, 2 real output (real part), 4 Ram-Lak
// ramp |j - N/2| / N.
template <class P>
__global__ void __launch_bounds__(P::NT)
ct_bluestein_kernel(const void* __restrict__ in, void* __restrict__ out,
                    const float2* __restrict__ chirp,
                    const float2* __restrict__ bhat,
                    const float2* __restrict__ tw, int N, long nrays,
                    int flags, float scale) {
  extern __shared__ float2 sm[];
  constexpr int L = P::L, E = P::E, T = P::T;
  constexpr int R1 = P::R1, RS = P::RS;
  const bool in_real = (flags & 1) != 0, out_real = (flags & 2) != 0;
  const bool ramp = (flags & 4) != 0;
  const int line = (int)(threadIdx.x / T);
  const int t = (int)(threadIdx.x % T);
  const long ray = (long)xcd_remap(blockIdx.x, gridDim.x) * P::LPW + line;
  const bool active = ray < nrays;
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int idx = t + i * T + r * (L / R1);
      float2 w = make_float2(0.f, 0.f);
      if (active && idx < N) {
        const long o = ray * N + idx;
        const float2 a = in_real
                             ? make_float2(static_cast<const float*>(in)[o], 0.f)
                             : static_cast<const float2*>(in)[o];
        w = cmul(a, chirp[idx]);
      }
      v[i * R1 + r] = w;
    }
  float2* ln = sm + line * P::LS;
  P::template run<-1>(v, ln, t, active, tw);
#pragma unroll
  for (int i = 0; i < E / RS; ++i)
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int k = t + i * T + r * (L / RS);
      v[i * RS + r] = cmul(v[i * RS + r], bhat[k]);
    }
  P::run_inv(v, ln, t, active, tw);
  if (!active) return;
  const int cidx = N / 2;
#pragma unroll
  for (int i = 0; i < E / R1; ++i)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int j = t + i * T + r * (L / R1);
      if (j < N) {
        const float2 q = cmul(v[i * R1 + r], chirp[j]);
        float f = scale;
        if (ramp) f *= (float)abs(j - cidx) / (float)N;
        const long o = ray * N + j;
        if (out_real)
          static_cast<float*>(out)[o] = q.x * f;
        else
          static_cast<float2*>(out)[o] = make_float2(q.x * f, q.y * f);
      }
    }
}

template <class P>
void launch_bluestein(const void* in, void* out, const float2* chirp,
                      const float2* bhat, int N, long nrays, int flags,
                      float scale, hipStream_t stream) {
  const size_t smem = sizeof(float2) * (size_t)P::LPW * P::LS;
  auto k = ct_bluestein_kernel<P>;
  allow_lds((const void*)k, smem);
  const unsigned nblk = (unsigned)cdiv(nrays, P::LPW);
  hipLaunchKernelGGL(k, dim3(nblk), dim3(P::NT), smem, stream, in, out, chirp,
                     bhat, fft_length(P::L).plan().tw, N, nrays, flags, scale);
  TC_HIP(hipGetLastError());
}

}  // namespace

bool bluestein_ct(const void* in, bool in_real, void* out, bool out_real,
                  const float2* chirp, const float2* bhat, int N, int L,
                  long nrays, bool ramp, float scale, hipStream_t stream) {
  if (nrays <= 0) return true;
  TC_REQUIRE(2 * N - 1 <= L, "Bluestein: length too short");
  const int flags = (in_real ? 1 : 0) | (out_real ? 2 : 0) | (ramp ? 4 : 0);
#define TC_BLUE(LL, PP)                                                    \
  case LL:                                                                 \
    launch_bluestein<PP>(in, out, chirp, bhat, N, nrays, flags, scale,     \
                         stream);                                          \
    return true;
  switch (L) {
    TC_BLUE(512, L512)
    TC_BLUE(768, L768)
    TC_BLUE(1024, L1024)
    TC_BLUE(1458, L1458)
    TC_BLUE(1536, L1536)
    TC_BLUE(2048, L2048)
    TC_BLUE(2304, L2304)
    TC_BLUE(2916, L2916)
    TC_BLUE(3072, L3072)
    TC_BLUE(4374, L4374)
    TC_BLUE(4608, L4608)
    TC_BLUE(5832, L5832)
    TC_BLUE(6144, L6144)
    TC_BLUE(8748, L8748)
    TC_BLUE(11664, L11664)
    default: return false;
  }
#undef TC_BLUE
}

bool fft_lines_ct_has(int L) {
  for (int l : kLineLengths)
    if (l == L) return true;
  return false;
}

int fft_lines_ct_length(int n) {
  int best = 0;
  for (int l : kLineLengths)
    if (l >= n && (best == 0 || l < best)) best = l;
  return best;
}

bool fft_lines_ct(const LineFFTArgs& a, int dir, hipStream_t stream) {
  switch (a.L) {
    case 512: return run_plan<L512>(a, dir, stream);
    case 768: return run_plan<L768>(a, dir, stream);
    case 1024: return run_plan<L1024>(a, dir, stream);
    case 1458: return run_plan<L1458>(a, dir, stream);
    case 1536: return run_plan<L1536>(a, dir, stream);
    case 2048: return run_plan<L2048>(a, dir, stream);
    case 2304: return run_plan<L2304>(a, dir, stream);
    case 2916: return run_plan<L2916>(a, dir, stream);
    case 3072: return run_plan<L3072>(a, dir, stream);
    case 4374: return run_plan<L4374>(a, dir, stream);
    case 4608: return run_plan<L4608>(a, dir, stream);
    case 5832: return run_plan<L5832>(a, dir, stream);
    case 6144: return run_plan<L6144>(a, dir, stream);
    case 8748: return run_plan<L8748>(a, dir, stream);
    case 11664: return run_plan<L11664>(a, dir, stream);
    default: return false;
  }
}

}  // namespace tc
