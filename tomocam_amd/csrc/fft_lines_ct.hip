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

template <int L_, int E_, int LPW_, class Rad>
struct LinePlan;
template <int L_, int E_, int LPW_, int... Rs>
struct LinePlan<L_, E_, LPW_, Radices<Rs...>> {
  static constexpr int L = L_, E = E_, LPW = LPW_, T = L_ / E_;
  static constexpr int NT = LPW_ * T;
  static constexpr int LS = lines_stride(ct::line_elems<L_, Rs...>(), LPW_);
  static constexpr int R1 = ct::RadixList<Rs...>::at(0);
  static constexpr int RS = ct::RadixList<Rs...>::at(sizeof...(Rs) - 1);
  template <int DIR>
  __device__ __forceinline__ static void run(float2* v, float2* line, int t,
                                             bool active,
                                             const float2* __restrict__ tw) {
    ct::transform<L_, E_, DIR, false, Rs...>(v, line, t, active, tw);
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
using L512 = LinePlan<512, 8, 4, Radices<8, 8, 8>>;
using L768 = LinePlan<768, 24, 8, Radices<24, 8, 4>>;
using L1024 = LinePlan<1024, 16, 4, Radices<16, 16, 4>>;
using L1458 = LinePlan<1458, 18, 3, Radices<18, 9, 9>>;
using L1536 = LinePlan<1536, 24, 4, Radices<24, 8, 8>>;
using L2048 = LinePlan<2048, 16, 2, Radices<16, 16, 8>>;
using L2304 = LinePlan<2304, 24, 2, Radices<24, 24, 4>>;
using L2916 = LinePlan<2916, 18, 2, Radices<18, 18, 9>>;
using L3072 = LinePlan<3072, 24, 2, Radices<24, 8, 4, 4>>;
using L4374 = LinePlan<4374, 18, 1, Radices<18, 9, 9, 3>>;
using L4608 = LinePlan<4608, 24, 1, Radices<24, 24, 8>>;
using L5832 = LinePlan<5832, 18, 1, Radices<18, 18, 18>>;
using L6144 = LinePlan<6144, 24, 1, Radices<24, 8, 8, 4>>;
using L8748 = LinePlan<8748, 18, 1, Radices<18, 18, 9, 3>>;
using L11664 = LinePlan<11664, 36, 1, Radices<36, 18, 18>>;

constexpr int kLineLengths[] = {512,  768,  1024, 1458, 1536,
                                2048, 2304, 2916, 3072, 4374,
                                4608, 5832, 6144, 8748, 11664};

}  // namespace

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
