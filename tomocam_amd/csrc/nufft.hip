// Polar-grid NUFFT (tile-owned LDS spreading, gather interpolation),
// Bluestein ray DFT and the projectors built on them.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "fft_device.h"
#include "fft_ops.h"
#include "lds_launch.h"
#include "nufft.h"

namespace tc {

// ---------------------------------------------------------------------------
// kernel function and host-side correction factors
// ---------------------------------------------------------------------------

// exp(beta (sqrt(1 - z^2) - 1)) on |z| <= 1, else 0.  Hardware sqrt and
// exp2 (about 1 ulp each): the kernel's own approximation error is ~1e-7.
__device__ __forceinline__ float es_kernel(float z, float beta) {
  const float s = 1.0f - z * z;
  const float e = __builtin_amdgcn_exp2f(
      beta * 1.44269504088896341f *
      (__builtin_amdgcn_sqrtf(fmaxf(s, 0.f)) - 1.0f));
  return s >= 0.f ? e : 0.f;
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton iteration on P_n).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) {
        p1 = 1.0;
        p2 = 0.0;
        for (int j = 1; j <= n; ++j) {
          double p3 = p2;
          p2 = p1;
          p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        break;
      }
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

// ---------------------------------------------------------------------------
// spreading (type 1): each workgroup owns a T x T tile of the fine grid and
// gathers every ray segment whose kernel footprint meets it (rays are
// enumerated analytically, so no sorting and no global atomics).
// ---------------------------------------------------------------------------

constexpr int SP_T = 32;
constexpr int SP_SB = 4;
constexpr int SP_LIST = 2048;

struct SpreadArgs {
  const float2* c;  // [nsl][P][N]
  float2* fine;     // [nsl][nf][nf]
  const double* ax;
  const double* ay;
  int P, N, cidx, nf, nsl, ntx;
  double umax;
  float beta;
};

// { t : lo < U0 + t a <= hi } intersected into [tlo, thi] (superset bounds)
__device__ __forceinline__ void ray_interval(double a, double U0, double lo,
                                             double hi, double& tlo,
                                             double& thi) {
  if (fabs(a) < 1e-12) {
    if (!(U0 > lo - 1e-9 && U0 <= hi + 1e-9)) {
      tlo = 1e30;
      thi = -1e30;
    }
    return;
  }
  double t1 = (lo - U0) / a, t2 = (hi - U0) / a;
  if (t1 > t2) {
    double q = t1;
    t1 = t2;
    t2 = q;
  }
  tlo = fmax(tlo, t1);
  thi = fmin(thi, t2);
}

// Inclusive prefix sum of one int per thread over a 256-thread block.
__device__ __forceinline__ int block_scan_incl(int v, int* wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  int off = 0;
  for (int w = 0; w < wid; ++w) off += wsum[w];
  __syncthreads();
  return v + off;
}

// Points per work item: a tile with more points is split into chunks whose
// partial tiles a fix-up pass adds.  Small grids (fewer than 4096 tiles, a
// few workgroups per CU) take 512 so the central tiles, where every ray
// passes, do not leave a tail; larger grids take 1024 (less fix-up).
// Measured with the longest-first order below: 8 x 512^2 at 360 angles
// 0.50 -> 0.44 ms with 512; 32 x 1454^2 at 720 angles 8.42 (1024) vs 8.75.
static int spread_chunk(int ntiles) { return ntiles < 4096 ? 512 : 1024; }

// Enumerates the ray segments (superset j ranges) of shift (U0, V0) whose
// kernel footprints meet the tile window; leaves seg_pj[k] = p | jlo << 11
// and exclusive offsets seg_off[0..nseg] in LDS and returns the number of
// points.  Called by all 256 threads (contains barriers).
__device__ int enumerate_segments(const SpreadArgs& a, double U0, double V0,
                                  double lox, double hix, double loy,
                                  double hiy, int* seg_pj, int* seg_off,
                                  int* wsum, int* nseg_sh, int* nsg_out) {
  const int tid = threadIdx.x;
  // Deterministic compaction: thread t examines angles t, t+256, ... and
  // writes its non-empty segments at a block-scanned offset, so every
  // workgroup of the same tile sees the same segment order.
  constexpr int MAXA = SP_LIST / 256;
  int jl[MAXA], jh[MAXA];
  int local = 0;
#pragma unroll
  for (int i = 0; i < MAXA; ++i) {
    const int p = tid + 256 * i;
    jl[i] = 1;
    jh[i] = 0;
    if (p < a.P) {
      double tlo = -1e30, thi = 1e30;
      ray_interval(a.ax[p], U0, lox, hix, tlo, thi);
      ray_interval(a.ay[p], V0, loy, hiy, tlo, thi);
      if (tlo <= thi) {
        tlo = fmax(tlo, -(double)a.N - 2.0);
        thi = fmin(thi, (double)a.N + 2.0);
        jl[i] = max(0, (int)floor(tlo) + a.cidx - 1);
        jh[i] = min(a.N - 1, (int)ceil(thi) + a.cidx + 1);
      }
    }
    local += jl[i] <= jh[i];
  }
  __syncthreads();  // previous users of the segment arrays are done
  const int incl0 = block_scan_incl(local, wsum);
  int pos = incl0 - local;
#pragma unroll
  for (int i = 0; i < MAXA; ++i) {
    if (jl[i] <= jh[i]) {
      seg_pj[pos] = (tid + 256 * i) | (jl[i] << 11);
      seg_off[pos] = jh[i] - jl[i] + 1;
      ++pos;
    }
  }
  if (tid == 255) *nseg_sh = incl0;
  __syncthreads();
  const int nsg = *nseg_sh;
  const int per = (nsg + 255) >> 8;
  const int start = tid * per;
  int lsum = 0;
  for (int k = 0; k < per; ++k)
    if (start + k < nsg) lsum += seg_off[start + k];
  const int incl = block_scan_incl(lsum, wsum);
  int run = incl - lsum;
  for (int k = 0; k < per; ++k) {
    const int idx = start + k;
    if (idx < nsg) {
      const int c = seg_off[idx];
      seg_off[idx] = run;
      run += c;
    }
  }
  if (tid == 255) seg_off[nsg] = incl;
  __syncthreads();
  *nsg_out = nsg;
  return seg_off[nsg];
}

struct TileWindow {
  int X0, Y0, X1, Y1;
  double lox, hix, loy, hiy;
};

__device__ __forceinline__ TileWindow tile_window(const SpreadArgs& a,
                                                  int tile) {
  TileWindow w;
  const int ty = tile / a.ntx;
  const int tx = tile - ty * a.ntx;
  w.X0 = tx * SP_T;
  w.Y0 = ty * SP_T;
  w.X1 = min(w.X0 + SP_T, a.nf);
  w.Y1 = min(w.Y0 + SP_T, a.nf);
  const double hw = 0.5 * NUFFT_W;
  w.lox = w.X0 - hw;
  w.hix = w.X1 - 1 + hw;
  w.loy = w.Y0 - hw;
  w.hiy = w.Y1 - 1 + hw;
  return w;
}

__device__ __forceinline__ bool shift_skipped(const SpreadArgs& a,
                                              const TileWindow& w, double U0,
                                              double V0) {
  return U0 + a.umax <= w.lox || U0 - a.umax > w.hix ||
         V0 + a.umax <= w.loy || V0 - a.umax > w.hiy;
}

// points per tile (all shifts), for building the work list
__global__ void __launch_bounds__(256) spread_count_kernel(SpreadArgs a,
                                                           int* counts) {
  __shared__ int seg_pj[SP_LIST];
  __shared__ int seg_off[SP_LIST + 1];
  __shared__ int wsum[4];
  __shared__ int nseg;
  const int tile = blockIdx.x;
  const TileWindow w = tile_window(a, tile);
  int tot = 0;
  for (int ky = -1; ky <= 1; ++ky)
    for (int kx = -1; kx <= 1; ++kx) {
      const double U0 = 0.5 * a.nf + kx * (double)a.nf;
      const double V0 = 0.5 * a.nf + ky * (double)a.nf;
      if (shift_skipped(a, w, U0, V0)) continue;
      int nsg;
      tot += enumerate_segments(a, U0, V0, w.lox, w.hix, w.loy, w.hiy, seg_pj,
                                seg_off, wsum, &nseg, &nsg);
    }
  if (threadIdx.x == 0) counts[tile] = tot;
}

// One workgroup per (work item, slice group): accumulates the points
// [f0, f1) of its tile in LDS and writes the tile to the fine grid (first
// chunk) or to a scratch slot (further chunks).  Points are taken in batches
// of SP_BATCH: each thread first prepares one point (cell origin, offsets of
// the first footprint cell, slice values, the 8-row bands of the tile its
// footprint meets) in LDS; then wave w adds the batch's points that meet
// band w (rows [8w, 8w + 8) of the tile, which it owns) with lane (a, b) on
// footprint cell (a, b).  Owned rows make the read-add-write race-free
// without atomics (LDS float atomics are ~30x slower than a plain
// read-add-write here), and each cell sums its points in list order, so the
// result is deterministic.  The segment list lives in dynamic LDS sized to
// the number of angles.
// Accumulator row stride (float2).  33 keeps a workgroup's LDS (41.5 KB
// static plus 8 bytes per angle) within a third of the CU's 160 KB up to
// 1639 angles, so three workgroups share a CU: at 1501 angles a 16 x 2908^2
// backprojection takes 36.2 instead of 43.2 ms with stride 36 (two
// workgroups per CU); at 720 and 360 angles, where stride 36 fits three as
// well, strides 33, 34 and 36 time the same.
constexpr int SP_RS = 33;
constexpr int SP_BATCH = 128;
static_assert(SP_SB == 4, "slice values are held as two float4 per point");

__global__ void __launch_bounds__(256)
spread_kernel(SpreadArgs a, const int4* __restrict__ work, float2* scratch) {
  __shared__ float2 acc[SP_SB][SP_T * SP_RS];
  __shared__ int4 pt_geo[SP_BATCH];      // l0x, l0y, bits(a0), bits(b0)
  __shared__ float4 pt_val[SP_BATCH][2];  // slices (0, 1), (2, 3)
  __shared__ int pt_band[SP_BATCH];
  __shared__ short wlist[4][SP_BATCH];
  __shared__ int wsum[4];
  __shared__ int nseg;
  extern __shared__ int seg_dyn[];  // seg_pj[P], seg_off[P + 1]
  int* seg_pj = seg_dyn;
  int* seg_off = seg_dyn + a.P;
  static_assert(SP_T == 32, "four 8-row bands per tile");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int qa = lane >> 3, qb = lane & 7;
  const int4 item = work[xcd_remap(blockIdx.x, gridDim.x)];
  const int s0 = blockIdx.y * SP_SB;
  const int ns = min(SP_SB, a.nsl - s0);
  const TileWindow w = tile_window(a, item.x);
  const int band0 = w.Y0 + 8 * wave;  // rows this wave owns
  const int band1 = min(band0 + 8, w.Y1);
  for (int i = tid; i < SP_SB * SP_T * SP_RS; i += 256)
    (&acc[0][0])[i] = make_float2(0.f, 0.f);
  const double hw = 0.5 * NUFFT_W;
  const long sstride = (long)a.P * a.N;
  const float zscale = 2.0f / NUFFT_W;
  int base = 0;
  for (int ky = -1; ky <= 1; ++ky) {
    for (int kx = -1; kx <= 1; ++kx) {
      const double U0 = 0.5 * a.nf + kx * (double)a.nf;
      const double V0 = 0.5 * a.nf + ky * (double)a.nf;
      if (shift_skipped(a, w, U0, V0)) continue;
      if (base >= item.z) continue;  // uniform: chunk already complete
      int nsg;
      const int total = enumerate_segments(a, U0, V0, w.lox, w.hix, w.loy,
                                           w.hiy, seg_pj, seg_off, wsum,
                                           &nseg, &nsg);
      const int fa = max(item.y, base) - base;
      const int fb = min(item.z, base + total) - base;
      for (int fb0 = fa; fb0 < fb; fb0 += SP_BATCH) {
        const int nb = min(SP_BATCH, fb - fb0);
        // prepare: thread t -> point fb0 + t
        if (tid < nb) {
          const int f = fb0 + tid;
          int lo = 0, hi = nsg;  // seg_off[lo] <= f < seg_off[hi]
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (seg_off[mid] <= f) lo = mid;
            else hi = mid;
          }
          const int pj = seg_pj[lo];
          const int p = pj & 2047;
          const int j = (pj >> 11) + (f - seg_off[lo]);
          const double t = (double)(j - a.cidx);
          const double u = U0 + t * a.ax[p];
          const double v = V0 + t * a.ay[p];
          const double cxu = ceil(u - hw), cyv = ceil(v - hw);
          const int l0x = (int)cxu, l0y = (int)cyv;
          pt_geo[tid] = make_int4(l0x, l0y,
                                  __builtin_bit_cast(int, (float)(cxu - u)),
                                  __builtin_bit_cast(int, (float)(cyv - v)));
          const float2* cp = a.c + (long)s0 * sstride + (long)p * a.N + j;
          float2 cv[SP_SB];
#pragma unroll
          for (int q = 0; q < SP_SB; ++q)
            cv[q] = q < ns ? cp[(long)q * sstride] : make_float2(0.f, 0.f);
          pt_val[tid][0] = make_float4(cv[0].x, cv[0].y, cv[1].x, cv[1].y);
          pt_val[tid][1] = make_float4(cv[2].x, cv[2].y, cv[3].x, cv[3].y);
          int bands = 0;
          if (l0x < w.X1 && l0x + NUFFT_W > w.X0) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              const int r0 = w.Y0 + 8 * b, r1 = min(r0 + 8, w.Y1);
              if (l0y < r1 && l0y + NUFFT_W > r0) bands |= 1 << b;
            }
          }
          pt_band[tid] = bands;
        }
        __syncthreads();
        // this wave's list of points meeting its band, kept in registers
        int cnt = 0;
        int lst[SP_BATCH / 64];
#pragma unroll
        for (int c = 0; c < SP_BATCH / 64; ++c) {
          const int idx = c * 64 + lane;
          const bool hit = idx < nb && ((pt_band[idx] >> wave) & 1);
          const unsigned long long bal = __ballot(hit);
          const int pos =
              cnt + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                              __builtin_amdgcn_mbcnt_lo(
                                                  (unsigned)bal, 0u));
          if (hit) wlist[wave][pos] = (short)idx;
          cnt += __popcll(bal);
        }
#pragma unroll
        for (int c = 0; c < SP_BATCH / 64; ++c) lst[c] = wlist[wave][c * 64 + lane];
        // software-pipelined: point k + 1's data is read from LDS while
        // point k's cells are updated
        auto entry = [&](int k) {
          return k < 64 ? __builtin_amdgcn_readlane(lst[0], k)
                        : __builtin_amdgcn_readlane(lst[1], k - 64);
        };
        static_assert(SP_BATCH == 128, "two list registers");
        int4 g = make_int4(0, 0, 0, 0);
        float4 v01 = make_float4(0.f, 0.f, 0.f, 0.f), v23 = v01;
        if (cnt > 0) {
          const int i = entry(0);
          g = pt_geo[i];
          v01 = pt_val[i][0];
          v23 = pt_val[i][1];
        }
        for (int k = 0; k < cnt; ++k) {
          const int in = entry(min(k + 1, cnt - 1));
          const int4 gn = pt_geo[in];
          const float4 v01n = pt_val[in][0];
          const float4 v23n = pt_val[in][1];
          const int cx = g.x + qb;
          const int cy = g.y + qa;
          // the point's 8 x weights (lanes 0-7) and 8 y weights (lanes
          // 8-15), one kernel evaluation per lane, fetched by lane (qa, qb)
          // from lanes qb and 8 + qa; every lane takes part, so this stays
          // outside the bounds test
          const int m = lane & 7;
          const float dl = __builtin_bit_cast(float, lane < 8 ? g.z : g.w) +
                           (float)m;
          const float wl = es_kernel(dl * zscale, a.beta);
          const float wx = __builtin_bit_cast(
              float, __builtin_amdgcn_ds_bpermute(
                         qb << 2, __builtin_bit_cast(int, wl)));
          const float wy = __builtin_bit_cast(
              float, __builtin_amdgcn_ds_bpermute(
                         (8 + qa) << 2, __builtin_bit_cast(int, wl)));
          if (cy >= band0 && cy < band1 && cx >= w.X0 && cx < w.X1) {
            const float wgt = wx * wy;
            const int cell = (cy - w.Y0) * SP_RS + (cx - w.X0);
            float2 a0 = acc[0][cell], a1 = acc[1][cell];
            float2 a2 = acc[2][cell], a3 = acc[3][cell];
            a0.x = fmaf(v01.x, wgt, a0.x);
            a0.y = fmaf(v01.y, wgt, a0.y);
            a1.x = fmaf(v01.z, wgt, a1.x);
            a1.y = fmaf(v01.w, wgt, a1.y);
            a2.x = fmaf(v23.x, wgt, a2.x);
            a2.y = fmaf(v23.y, wgt, a2.y);
            a3.x = fmaf(v23.z, wgt, a3.x);
            a3.y = fmaf(v23.w, wgt, a3.y);
            acc[0][cell] = a0;
            acc[1][cell] = a1;
            acc[2][cell] = a2;
            acc[3][cell] = a3;
          }
          g = gn;
          v01 = v01n;
          v23 = v23n;
        }
        __syncthreads();  // the batch buffers are rewritten next
      }
      base += total;
    }
  }
  __syncthreads();
  float2* dst = item.w < 0 ? nullptr
                           : scratch + ((long)item.w * gridDim.y + blockIdx.y) *
                                           (SP_SB * SP_T * SP_T);
  for (int i = tid; i < ns * SP_T * SP_T; i += 256) {
    const int s = i / (SP_T * SP_T);
    const int c = i - s * SP_T * SP_T;
    const int ly = c / SP_T, lx = c - ly * SP_T;
    const float2 val = acc[s][ly * SP_RS + lx];
    if (dst) {
      dst[i] = val;
    } else if (w.Y0 + ly < w.Y1 && w.X0 + lx < w.X1) {
      a.fine[((long)(s0 + s) * a.nf + w.Y0 + ly) * a.nf + w.X0 + lx] = val;
    }
  }
}

// Adds the scratch slots of multi-chunk tiles into the fine grid, in slot
// order (deterministic).
__global__ void __launch_bounds__(256)
spread_fixup_kernel(SpreadArgs a, const int4* __restrict__ fix,
                    const float2* __restrict__ scratch) {
  const int4 f = fix[blockIdx.x];
  const int s0 = blockIdx.y * SP_SB;
  const int ns = min(SP_SB, a.nsl - s0);
  const int ty = f.x / a.ntx, tx = f.x - (f.x / a.ntx) * a.ntx;
  const int X0 = tx * SP_T, Y0 = ty * SP_T;
  const int X1 = min(X0 + SP_T, a.nf), Y1 = min(Y0 + SP_T, a.nf);
  // blockIdx.z picks a quarter of the SP_SB x SP_T x SP_T elements, so a
  // tile's slots are summed by four workgroups; each thread sums its four
  // elements over the slots four slots at a time (16 loads in flight);
  // the per-element order is slot order
  constexpr int PART = SP_SB * SP_T * SP_T / 4;
  constexpr int PER = PART / 256;
  const int nel = ns * SP_T * SP_T;
  const int e0 = blockIdx.z * PART + threadIdx.x;
  const long slot_stride = (long)gridDim.y * (SP_SB * SP_T * SP_T);
  const float2* base = scratch + (long)f.y * slot_stride +
                       (long)blockIdx.y * (SP_SB * SP_T * SP_T);
  float2 sum[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) sum[e] = make_float2(0.f, 0.f);
  int k = 0;
  for (; k + 4 <= f.z; k += 4) {
    float2 q[4][PER];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        const int i = e0 + 256 * e;
        q[u][e] = i < nel ? base[(long)(k + u) * slot_stride + i]
                          : make_float2(0.f, 0.f);
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        sum[e].x += q[u][e].x;
        sum[e].y += q[u][e].y;
      }
  }
  for (; k < f.z; ++k) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int i = e0 + 256 * e;
      if (i < nel) {
        const float2 q = base[(long)k * slot_stride + i];
        sum[e].x += q.x;
        sum[e].y += q.y;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = e0 + 256 * e;
    const int s = i / (SP_T * SP_T);
    const int cell = i - s * SP_T * SP_T;
    const int ly = cell / SP_T, lx = cell - ly * SP_T;
    if (i < nel && Y0 + ly < Y1 && X0 + lx < X1) {
      float2* o = a.fine + ((long)(s0 + s) * a.nf + Y0 + ly) * a.nf + X0 + lx;
      const float2 cur = *o;
      *o = make_float2(cur.x + sum[e].x, cur.y + sum[e].y);
    }
  }
}

// ---------------------------------------------------------------------------
// interpolation (type 2): eight lanes per polar point, all slices of a batch
// ---------------------------------------------------------------------------

struct InterpArgs {
  const float2* fine;  // [nsl][nf][nf]
  float2* c;           // [nsl][P][N]
  const double* ax;
  const double* ay;
  int P, N, cidx, nf, nsl;
  float beta;
};

// Lane b of a point's group of eight reads footprint column b of every row
// (its eight rows from one slice are eight loads), so a wave's gather for
// one row covers eight points x 64 contiguous bytes instead of 64 scattered
// cells.  The lane weights its column sum by wx[b]; the eight columns are
// then summed across the group for eight slices at once by a transposing
// butterfly (exchange halves with lanes b^4, b^2, b^1: 14 shuffles per
// eight slices), after which lane b holds slice b of the chunk and writes
// it.  The summation order is fixed, so results are deterministic.
__device__ __forceinline__ float2 shfl_xor2(float2 v, int m) {
  return make_float2(__shfl_xor(v.x, m, 8), __shfl_xor(v.y, m, 8));
}

__global__ void __launch_bounds__(256) interp_kernel(InterpArgs a) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long pt = gid >> 3;
  const int b = (int)(gid & 7);
  // groups of eight are aligned to eight lanes: a group is all in or out
  if (pt >= (long)a.P * a.N) return;
  const int p = (int)(pt / a.N);
  const int j = (int)(pt - (long)p * a.N);
  const double hw = 0.5 * NUFFT_W;
  const double t = (double)(j - a.cidx);
  const double u = 0.5 * a.nf + t * a.ax[p];
  const double v = 0.5 * a.nf + t * a.ay[p];
  const int l0x = (int)ceil(u - hw);
  const int l0y = (int)ceil(v - hw);
  const float wxb = es_kernel((float)((l0x + b - u) * (2.0 / NUFFT_W)), a.beta);
  const float wyb = es_kernel((float)((l0y + b - v) * (2.0 / NUFFT_W)), a.beta);
  int qx = l0x + b;
  qx += qx < 0 ? a.nf : (qx >= a.nf ? -a.nf : 0);
  float wy[NUFFT_W];
  long roff[NUFFT_W];
#pragma unroll
  for (int r = 0; r < NUFFT_W; ++r) {
    wy[r] = __shfl(wyb, r, 8);
    int qy = l0y + r;
    qy += qy < 0 ? a.nf : (qy >= a.nf ? -a.nf : 0);
    roff[r] = (long)qy * a.nf + qx;
  }
  const long plane = (long)a.nf * a.nf;
  const long cplane = (long)a.P * a.N;
  for (int s0 = 0; s0 < a.nsl; s0 += 8) {
    float2 part[8];
#pragma unroll 2
    for (int q = 0; q < 8; ++q) {
      float2 acc = make_float2(0.f, 0.f);
      if (s0 + q < a.nsl) {
        const float2* g = a.fine + (long)(s0 + q) * plane;
        float2 val[NUFFT_W];
#pragma unroll
        for (int r = 0; r < NUFFT_W; ++r) val[r] = g[roff[r]];
#pragma unroll
        for (int r = 0; r < NUFFT_W; ++r) {
          acc.x = fmaf(val[r].x, wy[r], acc.x);
          acc.y = fmaf(val[r].y, wy[r], acc.y);
        }
      }
      part[q] = make_float2(acc.x * wxb, acc.y * wxb);
    }
    // lanes with bit m set keep the upper half of their slice set
    float2 h4[4], h2[2];
    const bool u4 = (b & 4) != 0, u2 = (b & 2) != 0, u1 = (b & 1) != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float2 keep = u4 ? part[4 + k] : part[k];
      const float2 give = u4 ? part[k] : part[4 + k];
      const float2 got = shfl_xor2(give, 4);
      h4[k] = make_float2(keep.x + got.x, keep.y + got.y);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const float2 keep = u2 ? h4[2 + k] : h4[k];
      const float2 give = u2 ? h4[k] : h4[2 + k];
      const float2 got = shfl_xor2(give, 2);
      h2[k] = make_float2(keep.x + got.x, keep.y + got.y);
    }
    const float2 keep = u1 ? h2[1] : h2[0];
    const float2 give = u1 ? h2[0] : h2[1];
    const float2 got = shfl_xor2(give, 1);
    // lane b now holds slice s0 + b (bits 4, 2, 1 of b chose the halves)
    if (s0 + b < a.nsl)
      a.c[(long)(s0 + b) * cplane + pt] =
          make_float2(keep.x + got.x, keep.y + got.y);
  }
}

// ---------------------------------------------------------------------------
// deconvolution: fine-grid spectrum <-> modes k in [-floor(nm/2), ceil(nm/2))
// ---------------------------------------------------------------------------

__global__ void deconv_real_kernel(const float2* __restrict__ fine, float* out,
                                   const float* __restrict__ corr, int nm,
                                   int nf, int nsl, float scale) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nm * nm;
  if (idx >= per * nsl) return;
  const int s = (int)(idx / per);
  const int rem = (int)(idx - (long)s * per);
  const int k2i = rem / nm, k1i = rem - (rem / nm) * nm;
  const int h = nm / 2;
  int i1 = k1i - h, i2 = k2i - h;
  i1 += i1 < 0 ? nf : 0;
  i2 += i2 < 0 ? nf : 0;
  const float2 v = fine[((long)s * nf + i2) * nf + i1];
  out[idx] = v.x * corr[k1i] * corr[k2i] * scale;
}

__global__ void precorrect_kernel(const float* __restrict__ f, float2* fine,
                                  const float* __restrict__ corr, int nm,
                                  int nf, int nsl) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nf * nf;
  if (idx >= per * nsl) return;
  const int s = (int)(idx / per);
  const long rem = idx - (long)s * per;
  const int i2 = (int)(rem / nf), i1 = (int)(rem - (long)i2 * nf);
  const int h = nm / 2, up = (nm + 1) / 2;
  int k1 = i1 < up ? i1 : (i1 >= nf - h ? i1 - nf : nf);
  int k2 = i2 < up ? i2 : (i2 >= nf - h ? i2 - nf : nf);
  float2 v = make_float2(0.f, 0.f);
  if (k1 != nf && k2 != nf) {
    const int k1i = k1 + h, k2i = k2 + h;
    v.x = f[((long)s * nm + k2i) * nm + k1i] * corr[k1i] * corr[k2i];
  }
  fine[idx] = v;
}

// ---------------------------------------------------------------------------
// PolarNufft
// ---------------------------------------------------------------------------

// Fine-grid size for at least m points: the smallest length with a
// compile-time line FFT plan when it exceeds the smallest even 2,3,5,7-smooth
// length by at most 1/8 (the planned transforms run several times faster
// than the runtime-plan kernel, and a larger grid only raises the
// upsampling factor above 2), else that smooth length.
static int fine_length(int m) {
  const int smooth = fft_next_smooth(m, true);
  const int c = fft_lines_ct_length(m);
  return (c > 0 && c <= smooth + smooth / 8) ? c : smooth;
}

PolarNufft::PolarNufft(const float* h_angles, int nproj, int npix, int nmodes)
    : P_(nproj), N_(npix), nm_(nmodes) {
  ArenaPause no_arena;  // cached plan, tables filled synchronously
  TC_REQUIRE(nproj >= 1 && npix >= 2 && nmodes >= 1, "NUFFT: bad geometry");
  TC_REQUIRE(nproj <= SP_LIST, "NUFFT: too many projection angles");
  TC_REQUIRE(npix < 65536, "NUFFT: too many samples per ray");
  nf_ = fine_length(std::max(2 * nmodes, 2 * NUFFT_W));
  beta_ = (float)(2.30 * NUFFT_W);
  std::vector<double> hx(P_), hy(P_);
  const double sc = (double)nf_ / (double)(npix - 1);
  for (int p = 0; p < P_; ++p) {
    const double th = (double)h_angles[p];
    hx[p] = std::cos(th) * sc;
    hy[p] = std::sin(th) * sc;
  }
  ax_.alloc(sizeof(double) * P_);
  ay_.alloc(sizeof(double) * P_);
  TC_HIP(hipMemcpy(ax_.as<void>(), hx.data(), sizeof(double) * P_,
                   hipMemcpyHostToDevice));
  TC_HIP(hipMemcpy(ay_.as<void>(), hy.data(), sizeof(double) * P_,
                   hipMemcpyHostToDevice));
  // corr(k) = (-1)^k h / psi_hat(k) = (-1)^k 2 / (w int phi(z) cos(k z w h/2))
  std::vector<double> zq, wq;
  gauss_legendre(120, zq, wq);
  const double h = 2.0 * M_PI / nf_;
  const double beta = (double)beta_;
  std::vector<float> corr(nm_);
  for (int ki = 0; ki < nm_; ++ki) {
    const int k = ki - nm_ / 2;
    double integ = 0.0;
    for (size_t q = 0; q < zq.size(); ++q) {
      const double z = zq[q];
      const double phi = std::exp(beta * (std::sqrt(1.0 - z * z) - 1.0));
      integ += wq[q] * phi * std::cos(k * z * NUFFT_W * h / 2.0);
    }
    double c = 2.0 / (NUFFT_W * integ);
    if (k & 1) c = -c;
    corr[ki] = (float)c;
  }
  corr_.alloc(sizeof(float) * nm_);
  TC_HIP(hipMemcpy(corr_.as<void>(), corr.data(), sizeof(float) * nm_,
                   hipMemcpyHostToDevice));
}

float2* PolarNufft::fine(int batch) {
  if (batch > fine_batch_) {
    fine_.alloc(sizeof(float2) * (size_t)batch * nf_ * nf_);
    fine_batch_ = batch;
  }
  return fine_.as<float2>();
}

// Points per tile of this geometry (a device count, read back once).
void PolarNufft::count_tiles(hipStream_t stream) {
  const int ntx = (int)cdiv(nf_, SP_T);
  const int ntiles = ntx * ntx;
  const int cidx = N_ / 2;
  SpreadArgs a{};
  a.ax = ax_.as<double>();
  a.ay = ay_.as<double>();
  a.P = P_;
  a.N = N_;
  a.cidx = cidx;
  a.nf = nf_;
  a.ntx = ntx;
  a.umax = std::max(cidx, N_ - 1 - cidx) * ((double)nf_ / (N_ - 1)) + 1.0;
  DevBuf dcounts(sizeof(int) * ntiles);
  hipLaunchKernelGGL(spread_count_kernel, dim3(ntiles), dim3(256), 0, stream,
                     a, dcounts.as<int>());
  TC_HIP(hipGetLastError());
  counts_.assign(ntiles, 0);
  TC_HIP(hipMemcpyAsync(counts_.data(), dcounts.as<void>(),
                        sizeof(int) * ntiles, hipMemcpyDeviceToHost, stream));
  TC_HIP(hipStreamSynchronize(stream));
}

// The work and fix-up lists for `chunk` points per item, uploaded to the
// plan's buffers (synchronous copies: the caller has no kernel reading the
// previous lists in flight).
void PolarNufft::make_work(int chunk) {
  ArenaPause no_arena;  // the plan's work lists, filled synchronously
  const int ntiles = (int)counts_.size();
  std::vector<int4> work, fix;
  int slot = 0;
  for (int t = 0; t < ntiles; ++t) {
    const int c = counts_[t];
    const int nch = std::max(1, (c + chunk - 1) / chunk);
    work.push_back(make_int4(t, 0, std::min(c, chunk), -1));
    if (nch > 1) {
      fix.push_back(make_int4(t, slot, nch - 1, 0));
      for (int k = 1; k < nch; ++k)
        work.push_back(make_int4(t, k * chunk, std::min(c, (k + 1) * chunk),
                                 slot++));
    }
  }
  // Heaviest items first: a workgroup's time grows with its points, and
  // workgroups start in block order, so the long ones must not be last
  // (16 x 2908^2 at 1501 angles: 28.6 -> 19.9 ms per backprojection).
  // Block b reads item xcd_remap(b) (each XCD takes a contiguous range of
  // the list), so rank b of the sorted order goes to index xcd_remap(b).
  // Each item writes its own tile or slot, and the fix-up adds slots in
  // slot order, so the result does not depend on the order.
  {
    std::vector<int4> byw(work);
    std::stable_sort(byw.begin(), byw.end(), [](const int4& x, const int4& y) {
      return x.z - x.y > y.z - y.y;
    });
    const int n = (int)byw.size();
    const int q = n >> 3, rmd = n & 7;
    for (int b = 0; b < n; ++b) {
      const int x = b & 7, t = b >> 3;
      work[x * q + (x < rmd ? x : rmd) + t] = byw[b];
    }
  }
  nwork_ = (int)work.size();
  nslots_ = slot;
  nfix_ = (int)fix.size();
  chunk_ = chunk;
  work_.alloc(sizeof(int4) * nwork_);
  TC_HIP(hipMemcpy(work_.as<void>(), work.data(), sizeof(int4) * nwork_,
                   hipMemcpyHostToDevice));
  if (nfix_) {
    fix_.alloc(sizeof(int4) * nfix_);
    TC_HIP(hipMemcpy(fix_.as<void>(), fix.data(), sizeof(int4) * nfix_,
                     hipMemcpyHostToDevice));
  }
}

// Spreading chunks picked by the autotuner, keyed by (device, angles, ray
// samples, fine grid, modes): later plans of the geometry, including those
// of concurrent batch workers (which do not tune), use the choice.
static std::mutex g_chunk_mu;
static std::map<std::tuple<int, int, int, int, int>, int> g_chunk_tuned;

static std::tuple<int, int, int, int, int> chunk_key(int P, int N, int nf,
                                                     int nm) {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  return std::make_tuple(dev, P, N, nf, nm);
}

// Slices per fine-grid batch: about 1 GB of fine grid, in whole groups of
// SP_SB slices (a spreading workgroup accumulates SP_SB slices; a partial
// group idles part of its accumulator and repeats the tile's segment
// enumeration and point set-up for fewer slices), rounding up while that
// stays within 1.25 GB.
static int fine_batch(int nf, int S) {
  const double per = 8.0 * nf * (double)nf, budget = 1.0e9;
  const int b0 = (int)std::max(1.0, std::floor(budget / per));
  int b = b0;
  const int up = (b0 + SP_SB - 1) / SP_SB * SP_SB;
  if (up * per <= 1.25 * budget)
    b = up;
  else if (b0 >= SP_SB)
    b = b0 / SP_SB * SP_SB;
  return std::max(1, std::min(S, b));
}

void PolarNufft::type1_real(const float2* c, float* out, int S, float scale,
                            hipStream_t stream) {
  TC_REQUIRE(S >= 1, "NUFFT: need at least one slice");
  const int B = fine_batch(nf_, S);
  float2* fg = fine(B);
  const int ntx = (int)cdiv(nf_, SP_T);
  const unsigned ngroups_max = cdiv(B, SP_SB);
  const int cidx = N_ / 2;
  const double umax =
      std::max(cidx, N_ - 1 - cidx) * ((double)nf_ / (N_ - 1)) + 1.0;
  auto args = [&](int s0, int nb) {
    SpreadArgs a;
    a.c = c + (long)s0 * P_ * N_;
    a.fine = fg;
    a.ax = ax_.as<double>();
    a.ay = ay_.as<double>();
    a.P = P_;
    a.N = N_;
    a.cidx = cidx;
    a.nf = nf_;
    a.nsl = nb;
    a.ntx = ntx;
    a.umax = umax;
    a.beta = beta_;
    return a;
  };
  const size_t seg_lds = sizeof(int) * (2 * (size_t)P_ + 1);
  // the fine grid of nb slices from the current work lists
  auto spread = [&](const SpreadArgs& a, int nb, float2* scratch) {
    const unsigned ng = cdiv(nb, SP_SB);
    hipLaunchKernelGGL(spread_kernel, dim3((unsigned)nwork_, ng), dim3(256),
                       seg_lds, stream, a, work_.as<int4>(), scratch);
    TC_HIP(hipGetLastError());
    if (nfix_) {
      hipLaunchKernelGGL(spread_fixup_kernel, dim3((unsigned)nfix_, ng, 4),
                         dim3(256), 0, stream, a, fix_.as<int4>(), scratch);
      TC_HIP(hipGetLastError());
    }
  };
  auto scratch_bytes = [&](unsigned ngroups) {
    return sizeof(float2) * (size_t)std::max(nslots_, 1) * ngroups * SP_SB *
           SP_T * SP_T;
  };
  if (!work_ready_) {
    count_tiles(stream);
    const int ntiles = (int)counts_.size();
    // Points per work item: TOMOCAM_HIP_SPREAD_CHUNK (>= 64), else the
    // autotuner's choice for this geometry, else (tuning allowed) the
    // fastest of the rule's choice, 512, 1024 and 2048 on this call's first
    // batch: best of two timed runs after a warm-up, the rule's choice kept
    // unless another is >= 5 % faster.  The chunk sets the order in which a
    // tile's points are summed, so the wide margin keeps near-ties (2048 is
    // within 0.7-1.3 % of 1024 on large grids) on the rule's choice and the
    // results the same from process to process.  Each work item writes its
    // own tile or slot, so a timing run only rewrites the fine grid.
    int chunk = 0;
    const char* e = std::getenv("TOMOCAM_HIP_SPREAD_CHUNK");
    if (e && std::atoi(e) >= 64) chunk = std::atoi(e);
    const auto key = chunk_key(P_, N_, nf_, nm_);
    if (!chunk) {
      std::lock_guard<std::mutex> lk(g_chunk_mu);
      auto it = g_chunk_tuned.find(key);
      if (it != g_chunk_tuned.end()) chunk = it->second;
    }
    if (!chunk && autotune_enabled()) {
      constexpr float kMargin = 0.05f;
      const int nb = std::min(B, S);
      std::vector<int> cands{spread_chunk(ntiles)};
      for (int k : {512, 1024, 2048})
        if (k != cands[0]) cands.push_back(k);
      hipEvent_t e0, e1;
      TC_HIP(hipEventCreate(&e0));
      TC_HIP(hipEventCreate(&e1));
      float best_ms = 0.f;
      try {
        for (int k : cands) {
          make_work(k);
          DevBuf scr(scratch_bytes(cdiv(nb, SP_SB)));
          float ms = 0.f;
          for (int rep = 0; rep < 3; ++rep) {
            TC_HIP(hipEventRecord(e0, stream));
            spread(args(0, nb), nb, scr.as<float2>());
            TC_HIP(hipEventRecord(e1, stream));
            TC_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            TC_HIP(hipEventElapsedTime(&t, e0, e1));
            if (rep == 1 || (rep > 1 && t < ms)) ms = t;
          }
          if (tune_log())
            std::fprintf(stderr,
                         "[tomocam] spread tune %d angles x %d, grid %d, %d "
                         "slices: chunk %d: %.3f ms\n", P_, N_, nf_, nb, k,
                         (double)ms);
          if (chunk == 0 || ms < (1.0f - kMargin) * best_ms) {
            chunk = k;
            best_ms = ms;
          }
        }
      } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      std::lock_guard<std::mutex> lk(g_chunk_mu);
      g_chunk_tuned[key] = chunk;
    }
    if (!chunk) chunk = spread_chunk(ntiles);
    if (chunk != chunk_) make_work(chunk);
    work_ready_ = true;
  }
  DevBuf scratch(scratch_bytes(ngroups_max));
  for (int s0 = 0; s0 < S; s0 += B) {
    const int nb = std::min(B, S - s0);
    const SpreadArgs a = args(s0, nb);
    spread(a, nb, scratch.as<float2>());
    // only the columns of modes [-nm/2, (nm+1)/2) are read back
    fft2d_pruned(fg, nb, nf_, (nm_ + 1) / 2, nm_ / 2, +1, true, stream);
    const long tot = (long)nb * nm_ * nm_;
    hipLaunchKernelGGL(deconv_real_kernel, dim3(cdiv(tot, 256)), dim3(256), 0,
                       stream, fg, out + (long)s0 * nm_ * nm_, corr_.as<float>(),
                       nm_, nf_, nb, scale);
    TC_HIP(hipGetLastError());
  }
  // the scratch slots are released on return: to the thread's arena
  // (reused in stream order) or by hipFree, which waits for the device
}

void PolarNufft::type2_real(const float* f, float2* c, int S,
                            hipStream_t stream) {
  TC_REQUIRE(S >= 1, "NUFFT: need at least one slice");
  const int B = fine_batch(nf_, S);
  float2* fg = fine(B);
  for (int s0 = 0; s0 < S; s0 += B) {
    const int nb = std::min(B, S - s0);
    const long tot = (long)nb * nf_ * nf_;
    hipLaunchKernelGGL(precorrect_kernel, dim3(cdiv(tot, 256)), dim3(256), 0,
                       stream, f + (long)s0 * nm_ * nm_, fg, corr_.as<float>(),
                       nm_, nf_, nb);
    TC_HIP(hipGetLastError());
    // only the columns of modes [-nm/2, (nm+1)/2) are non-zero
    fft2d_pruned(fg, nb, nf_, (nm_ + 1) / 2, nm_ / 2, -1, false, stream);
    InterpArgs a;
    a.fine = fg;
    a.c = c + (long)s0 * P_ * N_;
    a.ax = ax_.as<double>();
    a.ay = ay_.as<double>();
    a.P = P_;
    a.N = N_;
    a.cidx = N_ / 2;
    a.nf = nf_;
    a.nsl = nb;
    a.beta = beta_;
    hipLaunchKernelGGL(interp_kernel, dim3(cdiv(8L * P_ * N_, 256)),
                       dim3(256), 0, stream, a);
    TC_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// Bluestein ray DFT
// ---------------------------------------------------------------------------

template <int NB, bool IN_REAL, bool OUT_REAL, bool RAMP>
__global__ void __launch_bounds__(1024)
ray_dft_kernel(const void* __restrict__ in, void* __restrict__ out,
               const float2* __restrict__ chirp, const float2* __restrict__ bhat,
               FFTPlan P, int N, long nrays, int lpw, float scale) {
  extern __shared__ float2 sm[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int L = P.L;
  const long r0 = (long)blockIdx.x * lpw;
  for (int t = tid; t < lpw * L; t += nthr) {
    const int l = t / L;
    const int z = t - l * L;
    float2 v = make_float2(0.f, 0.f);
    if (z < N && r0 + l < nrays) {
      const long o = (r0 + l) * N + z;
      float2 a = IN_REAL ? make_float2(static_cast<const float*>(in)[o], 0.f)
                         : static_cast<const float2*>(in)[o];
      v = cmul(a, chirp[z]);
    }
    sm[t] = v;
  }
  __syncthreads();
  fft_lds<-1, NB>(sm, lpw, L, P, tid, nthr);
  for (int t = tid; t < lpw * L; t += nthr) {
    const int k = t - (t / L) * L;
    sm[t] = cmul(sm[t], bhat[k]);
  }
  __syncthreads();
  fft_lds<1, NB>(sm, lpw, L, P, tid, nthr);
  const int cidx = N / 2;
  for (int t = tid; t < lpw * N; t += nthr) {
    const int l = t / N;
    const int j = t - l * N;
    if (r0 + l >= nrays) continue;
    float2 v = cmul(sm[l * L + j], chirp[j]);
    float f = scale;
    if (RAMP) f *= (float)abs(j - cidx) / (float)N;
    const long o = (r0 + l) * N + j;
    if (OUT_REAL)
      static_cast<float*>(out)[o] = v.x * f;
    else
      static_cast<float2*>(out)[o] = cscale(v, f);
  }
}

RayDft::RayDft(int N, int sign) : N_(N) {
  ArenaPause no_arena;  // cached plan, tables filled synchronously
  TC_REQUIRE(N >= 1, "ray DFT: N must be >= 1");
  // the chirp-z length: a planned FFT length when within 1/8 of the
  // smallest 2,3,5,7-smooth one (the planned kernel runs forward transform,
  // multiply and inverse in registers)
  {
    const int smooth = fft_next_smooth(2 * N - 1, false);
    const int c = fft_lines_ct_length(2 * N - 1);
    L_ = (c > 0 && c <= smooth + smooth / 8) ? c : smooth;
  }
  const int c = N / 2;
  const long twoN = 2L * N;
  std::vector<float2> ch(N), b(L_, make_float2(0.f, 0.f));
  for (int z = 0; z < N; ++z) {
    const long m = z - c;
    const double ph = sign * M_PI * (double)((m * m) % twoN) / N;
    ch[z] = make_float2((float)std::cos(ph), (float)std::sin(ph));
  }
  for (long m = -(N - 1); m <= N - 1; ++m) {
    const double ph = -sign * M_PI * (double)((m * m) % twoN) / N;
    const long idx = m >= 0 ? m : L_ + m;
    b[idx] = make_float2((float)(std::cos(ph) / L_), (float)(std::sin(ph) / L_));
  }
  chirp_.alloc(sizeof(float2) * N);
  bhat_.alloc(sizeof(float2) * L_);
  TC_HIP(hipMemcpy(chirp_.as<void>(), ch.data(), sizeof(float2) * N,
                   hipMemcpyHostToDevice));
  TC_HIP(hipMemcpy(bhat_.as<void>(), b.data(), sizeof(float2) * L_,
                   hipMemcpyHostToDevice));
  LineFFTArgs a;
  a.in = bhat_.as<float2>();
  a.out = bhat_.as<float2>();
  a.L = L_;
  a.lines_per_batch = 1;
  a.nbatch = 1;
  a.lstride = L_;
  a.bstride = L_;
  a.nin = L_;
  a.nout = L_;
  fft_lines(a, -1, nullptr);
  TC_HIP(hipDeviceSynchronize());
}

void RayDft::apply(const void* in, bool in_real, void* out, bool out_real,
                   long nrays, bool ramp, float scale, hipStream_t stream) {
  if (nrays <= 0) return;
  {
    const char* e = std::getenv("TOMOCAM_HIP_CT_LINES");
    if (!(e && e[0] == '0') &&
        bluestein_ct(in, in_real, out, out_real, chirp_.as<float2>(),
                     bhat_.as<float2>(), N_, L_, nrays, ramp, scale, stream))
      return;
  }
  const FFTLength& F = fft_length(L_);
  LdsLaunch ll = plan_lds_launch(F, 4, 80 * 1024);
  unsigned nblk = cdiv(nrays, ll.lpw);
#define TC_RAY(IR, OR, RP)                                                   \
  if (in_real == IR && out_real == OR && ramp == RP) {                       \
    TC_DISPATCH_NB(ll.nb, {                                                  \
      auto k = ray_dft_kernel<NB, IR, OR, RP>;                               \
      allow_lds((const void*)k, ll.smem);                                    \
      hipLaunchKernelGGL(k, dim3(nblk), dim3(ll.nthr), ll.smem, stream, in,  \
                         out, chirp_.as<float2>(), bhat_.as<float2>(),       \
                         F.plan(), N_, nrays, ll.lpw, scale);                \
    });                                                                      \
    TC_HIP(hipGetLastError());                                               \
    return;                                                                  \
  }
  TC_RAY(true, false, false)
  TC_RAY(true, false, true)
  TC_RAY(false, true, false)
  TC_RAY(false, false, false)
  TC_RAY(true, true, false)
#undef TC_RAY
  throw Error(TC_ERR_INTERNAL, "ray DFT: unsupported combination");
}

// ---------------------------------------------------------------------------
// plan caches
// ---------------------------------------------------------------------------

// A PolarNufft keeps scratch (the fine grid) between calls, so threads that
// run concurrently use plans of their own: the cache key holds a plan slot,
// taken by each thread on its first plan request from a pool (lowest free
// number) and given back when the thread ends.  Threads that follow one
// another, such as the batch workers of successive calls, reuse the slots
// and with them the cached plans.
static std::mutex g_slot_mu;
static std::vector<char> g_slot_used;

struct PlanSlot {
  int id = -1;
  ~PlanSlot() {
    if (id >= 0) {
      std::lock_guard<std::mutex> lk(g_slot_mu);
      g_slot_used[id] = 0;
    }
  }
};
static thread_local PlanSlot t_slot;

static int plan_slot() {
  if (t_slot.id < 0) {
    std::lock_guard<std::mutex> lk(g_slot_mu);
    int i = 0;
    while (i < (int)g_slot_used.size() && g_slot_used[i]) ++i;
    if (i == (int)g_slot_used.size()) g_slot_used.push_back(0);
    g_slot_used[i] = 1;
    t_slot.id = i;
  }
  return t_slot.id;
}

PolarNufft& nufft_plan(const float* h_angles, int nproj, int npix,
                       int nmodes) {
  static std::mutex mu;
  using Key = std::tuple<int, int, int, int, int, std::vector<float>>;
  static std::map<Key, std::unique_ptr<PolarNufft>> cache;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  std::vector<float> ang(h_angles, h_angles + nproj);
  const int slot = plan_slot();
  Key key(slot, dev, nproj, npix, nmodes, ang);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    // at most 16 plans per slot: drop this slot's (only this thread uses
    // them, and it holds none while it asks for a plan)
    int mine = 0;
    for (auto& kv : cache) mine += std::get<0>(kv.first) == slot;
    if (mine >= 16)
      for (auto i = cache.begin(); i != cache.end();)
        i = std::get<0>(i->first) == slot ? cache.erase(i) : std::next(i);
    it = cache
             .emplace(key, std::make_unique<PolarNufft>(h_angles, nproj, npix,
                                                        nmodes))
             .first;
  }
  return *it->second;
}

RayDft& ray_dft_plan(int N, int sign) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, std::unique_ptr<RayDft>> cache;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, N, sign);
  auto it = cache.find(key);
  if (it == cache.end())
    it = cache.emplace(key, std::make_unique<RayDft>(N, sign)).first;
  return *it->second;
}

// ---------------------------------------------------------------------------
// projectors
// ---------------------------------------------------------------------------

void project(const float* img, float* sino, int S, int N, const float* h_angles,
             int P, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && N >= 2 && P >= 1, "project: bad shape");
  PolarNufft& nu = nufft_plan(h_angles, P, N, N);
  RayDft& rd = ray_dft_plan(N, +1);
  DevBuf c(sizeof(float2) * (size_t)S * P * N);
  nu.type2_real(img, c.as<float2>(), S, stream);
  rd.apply(c.as<float2>(), false, sino, true, (long)S * P, false,
           1.0f / ((float)N * (float)N), stream);
}

void backproject(const float* sino, float* img, int S, int P, int N,
                 const float* h_angles, bool fbp, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && N >= 2 && P >= 1, "backproject: bad shape");
  PolarNufft& nu = nufft_plan(h_angles, P, N, N);
  RayDft& rd = ray_dft_plan(N, -1);
  DevBuf c(sizeof(float2) * (size_t)S * P * N);
  rd.apply(sino, true, c.as<float2>(), false, (long)S * P, fbp, 1.0f, stream);
  nu.type1_real(c.as<float2>(), img, S, 1.0f / ((float)N * (float)N), stream);
}

__global__ void fill_ones_kernel(float2* c, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = make_float2(1.f, 0.f);
}

void polar_psf(const float* h_angles, int P, int n, float* d_psf,
               hipStream_t stream) {
  TC_REQUIRE(P >= 1 && n >= 2, "PSF: bad geometry");
  PolarNufft& nu = nufft_plan(h_angles, P, n, 2 * n - 1);
  DevBuf ones(sizeof(float2) * (size_t)P * n);
  const long tot = (long)P * n;
  hipLaunchKernelGGL(fill_ones_kernel, dim3(cdiv(tot, 256)), dim3(256), 0,
                     stream, ones.as<float2>(), tot);
  TC_HIP(hipGetLastError());
  nu.type1_real(ones.as<float2>(), d_psf, 1, 1.0f, stream);
  TC_HIP(hipStreamSynchronize(stream));
}

}  // namespace tc
