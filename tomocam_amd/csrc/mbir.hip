// qGGMRF stencil, fused MBIR gradient/commit kernels and the run2 driver.
#include <hip/hip_fp16.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lds_launch.h"
#include "mbir.h"

namespace tc {

// ---------------------------------------------------------------------------
// qGGMRF potential derivative (q = 2, c = 1e-3)
//   rho'(d) = d / sigma^2 / (c + t) * (2 - (2 - p) t / (c + t)),
//   t = (|d| / sigma)^(2 - p)
// Reference: d_pot_func, src/gpu/potential_function.cuh:90-108.
// ---------------------------------------------------------------------------

constexpr float MRF_C = 0.001f;
// 26-neighbour weights (src/gpu/potential_function.cuh:38-51): faces, edges,
// corners; centre weight is 0.
constexpr float W_FACE = 0.0523f;
constexpr float W_EDGE = 0.0370f;
constexpr float W_CORNER = 0.0302f;

// two FP16 values in one word (a in the low half)
__device__ __forceinline__ unsigned pack_h2(__half a, __half b) {
  return (unsigned)__half_as_ushort(a) |
         ((unsigned)__half_as_ushort(b) << 16);
}

struct RhoParams {
  float inv_sigma;
  float inv_sigma2;
  float qmp;   // q - p
  float klog;  // (q - p) log2(1 / sigma)
};

// A neighbour class's constants for the stencil's rho' (ws2 = w / sigma^2):
// a = ws2 (q-p) c, b = ws2 (2 - (q-p))
struct RhoW {
  float a, b;
};

// rho'(d) * w for a neighbour of weight w (ws2 = w / sigma^2):
//   ws2 d r (2 - (q-p) t r),  t = (|d|/sigma)^(q-p), r = 1/(c + t),
// with t r = 1 - c r:  d r (W.b + W.a r), W.a = ws2 (q-p) c,
// W.b = ws2 (2 - (q-p)) (a sum of non-negative terms: no cancellation), and
// t = 2^((q-p) log2|d| + (q-p) log2(1/sigma)): nine VALU operations per
// pair term; log2(0) = -inf gives t = 0
__device__ __forceinline__ float wdrho(float d, RhoW W, const RhoParams& P) {
  const float t = __builtin_amdgcn_exp2f(
      fmaf(P.qmp, __builtin_amdgcn_logf(fabsf(d)), P.klog));
  const float r = __builtin_amdgcn_rcpf(MRF_C + t);
  return d * (r * fmaf(W.a, r, W.b));
}

struct QArgs {
  const float* sol;    // y = sol (+ beta (sol - xprev) in MBIR mode)
  const float* xprev;
  const float* Tsol;
  const float* Tx;
  const float* b;
  float* g;
  double* partials;    // NDOT per workgroup (MBIR mode)
  // FP16 gradient (mixed precision): g16 = g * gsc[slice] in the row-pair
  // layout (half_pair_bytes), and the largest |g| of each wave and plane
  // into gmaxp
  unsigned* g16;
  const float* gsc;
  float* gmaxp;
  int S, R, C;
  int zchunk;
  float beta;
  RhoParams rp;
  RhoW w_face, w_edge, w_corner;
};

// Pair-symmetric stencil.  rho' is odd, so the term a voxel a receives
// from neighbour a + o is minus the term a + o receives from a:
//   g(a) = sum_{o in F} phi_o(a) - sum_{o in F} phi_o(a - o),
//   phi_o(p) = w_o rho'(V(p) - V(p + o)),
// with F the 13 "forward" offsets (dz = +1: 9; dz = 0, dy = +1: 3;
// dz = dy = 0, dx = +1: 1) and V the volume extended by replication.  The
// extension makes the identity exact at the faces too: for a in the volume
// and a - o outside, w rho'(v_a - V(a - o)) = -phi_o(a - o) because
// V(a - o + o) = v_a, which is the reference's clamped-neighbour rule.
//
// Layout: one wave holds a strip of QROWS output rows and QCOLS = 62 output
// columns; lane l is column c0 - 1 + l (lanes 0 and 63 are the x halo) and
// each lane keeps its column's QROWS + 2 rows (y halo) of two planes in
// registers.  x neighbours come from the adjacent lane (DPP wave shift), y
// neighbours from the lane's own registers, and the wave marches in z,
// carrying the backward terms of the dz = +1 pairs to the next plane.
// Each pair is evaluated once: 13 QROWS + 9 rho' per strip row-plane
// instead of 26 QROWS.
constexpr int QLANES = 64;
constexpr int QCOLS = QLANES - 2;
#ifndef TC_QROWS
#define TC_QROWS 8
#endif
constexpr int QROWS = TC_QROWS;
constexpr int QWAVES = 4;  // strips per workgroup, stacked in y
constexpr int QJ = QROWS + 2;
// line-search dot products per workgroup (MBIR mode):
//   y.Ty, y.b, g.(Ty - b), g.g, d.d, d.g   (d = sol - x)
constexpr int NDOT = 6;

// value held by lane - 1 (column x - 1) / lane + 1 (column x + 1); the end
// lanes receive 0 and only feed halo lanes, whose results are discarded.
// Plain DPP moves (bound_ctrl, full masks) fold into the consuming VALU op.
__device__ __forceinline__ float from_left(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                      0x138 /* wave_shr:1 */, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                      0x130 /* wave_shl:1 */, 0xf, 0xf, true));
}

template <int DX>
__device__ __forceinline__ float shift_x(float v) {
  if constexpr (DX > 0) return from_left(v);
  else if constexpr (DX < 0) return from_right(v);
  else return v;
}

// The pair terms of forward offset (DZ, DY, DX) between plane `lo` and
// plane `hi` (= lo for DZ = 0).  FWD: add phi(a) to glo for the strip's
// rows; always subtract phi(a - o) from ghi.
template <int DY, int DX, bool FWD>
__device__ __forceinline__ void qpair(const float (&lo)[QJ],
                                      const float (&hi)[QJ], float (&glo)[QJ],
                                      float (&ghi)[QJ], RhoW W,
                                      const RhoParams& P) {
  constexpr int jlo = FWD ? (DY > 0 ? 0 : 1) : 1 - DY;
  constexpr int jhi = FWD ? (DY < 0 ? QROWS + 1 : QROWS) : QROWS - DY;
  float phi[QJ];
#pragma unroll
  for (int j = jlo; j <= jhi; ++j) {
    // V(a + o): column x + DX of plane hi
    phi[j] = wdrho(lo[j] - shift_x<-DX>(hi[j + DY]), W, P);
  }
  if constexpr (FWD) {
#pragma unroll
    for (int j = 1; j <= QROWS; ++j) glo[j] += phi[j];
  }
#pragma unroll
  for (int j = 1; j <= QROWS; ++j) ghi[j] -= shift_x<DX>(phi[j - DY]);
  // Materialise the accumulators after each pair set.  Without this the
  // compiler defers the additions (to the output branch or the next plane)
  // and keeps every term live: > 256 VGPRs.  8-9 independent rho' chains
  // per set are enough to hide the latency.
#pragma unroll
  for (int j = 1; j <= QROWS; ++j) {
    if constexpr (FWD) asm volatile("" : "+v"(glo[j]));
    asm volatile("" : "+v"(ghi[j]));
  }
}

// the 9 dz = +1 pairs between planes lo and hi
template <bool FWD>
__device__ __forceinline__ void qcross(const float (&lo)[QJ],
                                       const float (&hi)[QJ],
                                       float (&glo)[QJ], float (&ghi)[QJ],
                                       const QArgs& a) {
  qpair<0, 0, FWD>(lo, hi, glo, ghi, a.w_face, a.rp);
  qpair<1, 0, FWD>(lo, hi, glo, ghi, a.w_edge, a.rp);
  qpair<-1, 0, FWD>(lo, hi, glo, ghi, a.w_edge, a.rp);
  qpair<0, 1, FWD>(lo, hi, glo, ghi, a.w_edge, a.rp);
  qpair<0, -1, FWD>(lo, hi, glo, ghi, a.w_edge, a.rp);
  qpair<1, 1, FWD>(lo, hi, glo, ghi, a.w_corner, a.rp);
  qpair<1, -1, FWD>(lo, hi, glo, ghi, a.w_corner, a.rp);
  qpair<-1, 1, FWD>(lo, hi, glo, ghi, a.w_corner, a.rp);
  qpair<-1, -1, FWD>(lo, hi, glo, ghi, a.w_corner, a.rp);
}

// the 4 in-plane forward pairs: (0,0,+1), (0,+1,-1), (0,+1,0), (0,+1,+1)
__device__ __forceinline__ void qplane(const float (&v)[QJ], float (&g)[QJ],
                                       const QArgs& a) {
  qpair<0, 1, true>(v, v, g, g, a.w_face, a.rp);
  qpair<1, 0, true>(v, v, g, g, a.w_face, a.rp);
  qpair<1, 1, true>(v, v, g, g, a.w_edge, a.rp);
  qpair<1, -1, true>(v, v, g, g, a.w_edge, a.rp);
}

// y (= sol, or sol + beta (sol - xprev)) of plane z at the strip's
// in-plane offsets ro (rows r0 - 1 .. r0 + QROWS, replicated at the faces).
// Plane pointers are wave-uniform and offsets 32-bit, so every load is
// scalar base + one VGPR.
template <typename T>
__device__ __forceinline__ T& at_byte(T* base, unsigned off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T& at_byte(const T* base, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     off);
}

template <bool HAS_X>
__device__ __forceinline__ void qload(float (&v)[QJ], const QArgs& a, int z,
                                      const unsigned (&ro)[QJ]) {
  const long zo = (long)min(max(z, 0), a.S - 1) * a.R * a.C;
  const float* sp = a.sol + zo;
  const float* xp = a.xprev + zo;
#pragma unroll
  for (int j = 0; j < QJ; ++j) {
    float s = at_byte(sp, ro[j]);
    if constexpr (HAS_X) s = fmaf(a.beta, s - at_byte(xp, ro[j]), s);
    v[j] = s;
  }
}

// MODE 0: g += R'(x).  MODE 1 (MBIR): y = sol + beta (sol - xprev),
// g = Tsol + beta (Tsol - Tx) - b + R'(y), plus the line-search dots.
// HAS_X = false means xprev == sol and Tx == Tsol (y = sol, d = 0).
// G16 (MODE 1): g is stored in FP16 times the plane's scale (saturating at
// the largest finite value), and the dots use the stored value.
template <int MODE, bool HAS_X, bool G16 = false>
__global__ void __launch_bounds__(256) qggmrf_kernel(QArgs a) {
  __shared__ double red[16];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ntx = (a.C + QCOLS - 1) / QCOLS;
  const int nty = (a.R + QROWS * QWAVES - 1) / (QROWS * QWAVES);
  const int ntiles = ntx * nty;
  const int bl = xcd_remap(blockIdx.x, gridDim.x);
  const int zc = bl / ntiles;
  const int tile = bl - zc * ntiles;
  const int ty = tile / ntx;
  const int tx = tile - ty * ntx;
  const int x = tx * QCOLS - 1 + lane;
  const int xc = min(max(x, 0), a.C - 1);
  const int r0 = (ty * QWAVES + wave) * QROWS;
  const int z0 = zc * a.zchunk;
  const int z1 = min(a.S, z0 + a.zchunk);
  const bool xout = lane >= 1 && lane <= QCOLS && x < a.C;
  const long plane = (long)a.R * a.C;
  // in-plane byte offsets of the strip's rows (planes hold < 2^30
  // elements); for output rows (r < R) with x < C they are the outputs'
  // offsets too
  unsigned ro[QJ];
#pragma unroll
  for (int j = 0; j < QJ; ++j)
    ro[j] = 4u * (unsigned)(min(max(r0 - 1 + j, 0), a.R - 1) * a.C + xc);

  float vc[QJ], vn[QJ];
  float gc[QJ], gn[QJ];
#pragma unroll
  for (int j = 0; j < QJ; ++j) gc[j] = 0.f;

  // backward terms of plane z0 from the pairs (z0 - 1, z0)
  qload<HAS_X>(vn, a, z0 - 1, ro);
  qload<HAS_X>(vc, a, z0, ro);
  qcross<false>(vn, vc, gn, gc, a);

  double acc[NDOT] = {0, 0, 0, 0, 0, 0};
  for (int z = z0; z < z1; ++z) {
    // opaque per iteration, so the 32-bit offsets are not hoisted as 64-bit
    // values and every access uses the scalar-base + VGPR-offset form
#pragma unroll
    for (int j = 0; j < QJ; ++j) asm volatile("" : "+v"(ro[j]));
    qload<HAS_X>(vn, a, z + 1, ro);
#pragma unroll
    for (int j = 0; j < QJ; ++j) gn[j] = 0.f;
    qplane(vc, gc, a);
    qcross<true>(vc, vn, gc, gn, a);

    float gm = 0.f;  // largest |g| of the wave in this plane (G16)
    if (xout) {
      const long zo = (long)z * plane;
      float* gp = a.g + zo;
      // FP16 g in the row-pair layout: plane z, pair q = row / 2, column x
      unsigned* g16w =
          G16 ? a.g16 + ((long)z * ((a.R + 1) / 2) + r0 / 2) * a.C + x
              : nullptr;
      __half hprev = __float2half_rn(0.f);
      float gs = 1.f, gsi = 1.f;
      if constexpr (G16) {
        gs = a.gsc[z];
        gsi = 1.0f / gs;
      }
      const float* bp = a.b + zo;
      const float* tsp = a.Tsol + zo;
      const float* txp = a.Tx + zo;
      const float* sp = a.sol + zo;
      const float* xpp = a.xprev + zo;
      // Every operand of the strip's rows is loaded before the first
      // store: a store between two loads orders them (the pointers may
      // alias for all the compiler knows), and loads inside the rows'
      // bounds checks cannot be hoisted, so each would wait out its
      // latency in turn.  The offsets are clamped into the plane, so rows
      // past the last one read in-bounds values that go unused.
      constexpr int QE = MODE == 1 ? QROWS + 1 : 1;
      float bv[QE], tv[QE], dv[MODE == 1 && HAS_X ? QROWS + 1 : 1];
      if constexpr (MODE == 1) {
#pragma unroll
        for (int j = 1; j <= QROWS; ++j) {
          const unsigned o = ro[j];
          bv[j] = at_byte(bp, o);
          tv[j] = at_byte(tsp, o);
          if constexpr (HAS_X) {
            tv[j] = fmaf(a.beta, tv[j] - at_byte(txp, o), tv[j]);
            dv[j] = at_byte(sp, o) - at_byte(xpp, o);
          }
        }
      }
#pragma unroll
      for (int j = 1; j <= QROWS; ++j) {
        if (r0 + j - 1 < a.R) {
          const unsigned o = ro[j];
          if constexpr (MODE == 0) {
            at_byte(gp, o) += gc[j];
          } else {
            const float v = vc[j];
            const float bb = bv[j];
            const float tyv = tv[j];
            float gv = (tyv - bb) + gc[j];
            if constexpr (G16) {
              gm = fmaxf(gm, fabsf(gv));
              const __half h =
                  __float2half_rn(fminf(fmaxf(gv * gs, -65504.f), 65504.f));
              // r0 is even: rows r0 + j - 1 pair up as (odd j, even j); a
              // last row of even index goes out with a zero partner
              if ((j & 1) == 0)
                g16w[(long)(j / 2 - 1) * a.C] = pack_h2(hprev, h);
              else if (r0 + j == a.R)
                g16w[(long)(j / 2) * a.C] = pack_h2(h, __float2half_rn(0.f));
              hprev = h;
              gv = __half2float(h) * gsi;
            } else {
              at_byte(gp, o) = gv;
            }
            acc[0] += (double)v * tyv;
            acc[1] += (double)v * bb;
            acc[2] += (double)gv * (tyv - bb);
            acc[3] += (double)gv * gv;
            if constexpr (HAS_X) {
              const float d = dv[j];
              acc[4] += (double)d * d;
              acc[5] += (double)d * gv;
            }
          }
        }
      }
    }
    if constexpr (G16) {
      for (int o = 32; o > 0; o >>= 1) gm = fmaxf(gm, __shfl_xor(gm, o));
      if (lane == 0)
        a.gmaxp[((long)bl * QWAVES + wave) * a.zchunk + (z - z0)] = gm;
    }
#pragma unroll
    for (int j = 0; j < QJ; ++j) {
      vc[j] = vn[j];
      gc[j] = gn[j];
    }
  }
  if (MODE == 1) {
#pragma unroll
    for (int k = 0; k < NDOT; ++k) {
      double tot = block_sum(acc[k], red);
      if (threadIdx.x == 0) a.partials[(long)bl * NDOT + k] = tot;
    }
  }
}

static RhoParams make_rho(float sigma, float p) {
  TC_REQUIRE(sigma > 0.f && std::isfinite(sigma), "qGGMRF: sigma must be > 0");
  RhoParams r;
  r.inv_sigma = 1.0f / sigma;
  r.inv_sigma2 = 1.0f / (sigma * sigma);
  r.qmp = 2.0f - p;
  r.klog = (float)((2.0 - (double)p) * std::log2(1.0 / (double)sigma));
  return r;
}

static void set_rho(QArgs& a, float sigma, float p) {
  a.rp = make_rho(sigma, p);
  const float qmp = a.rp.qmp;
  auto cls = [&](float w) {
    const float ws2 = w * a.rp.inv_sigma2;
    return RhoW{ws2 * qmp * MRF_C, ws2 * (2.0f - qmp)};
  };
  a.w_face = cls(W_FACE);
  a.w_edge = cls(W_EDGE);
  a.w_corner = cls(W_CORNER);
}

// z chunking: a chunk costs its planes plus ~0.7 plane for the pairs with
// the plane before it; waves are spread over the SIMDs, so pick the chunk
// that minimises (rounds of waves per SIMD) x (planes per wave).
static void qggmrf_grid(int S, int R, int C, int* zchunk, unsigned* nblk) {
  static int simds = 0;
  if (!simds) {
    int dev = 0, cus = 0;
    TC_HIP(hipGetDevice(&dev));
    TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                 dev));
    simds = 4 * (cus > 0 ? cus : 256);
  }
  const long ntiles = cdiv(C, QCOLS) * cdiv(R, QROWS * QWAVES);
  int best = S;
  double best_cost = 1e300;
  for (int nz = 1; nz <= S; ++nz) {
    const int zc = (int)cdiv(S, nz);  // shortest chunk giving nz chunks
    if (cdiv(S, zc) != nz) continue;
    const long waves = ntiles * QWAVES * nz;
    const double rounds = (double)cdiv(waves, simds);
    const double cost = rounds * (zc + 0.7);
    if (cost < best_cost * 0.999) {
      best_cost = cost;
      best = zc;
    }
  }
  // TOMOCAM_HIP_STENCIL_ZCHUNK overrides the planes per wave
  const char* e = std::getenv("TOMOCAM_HIP_STENCIL_ZCHUNK");
  if (e && std::atoi(e) >= 1) best = std::min(std::atoi(e), S);
  *zchunk = best;
  *nblk = (unsigned)(ntiles * cdiv(S, best));
}

void add_total_var(const float* x, float* g, int S, int R, int C, float sigma,
                   float p, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1, "add_total_var: empty volume");
  QArgs a{};
  a.sol = x;
  a.g = g;
  a.S = S;
  a.R = R;
  a.C = C;
  set_rho(a, sigma, p);
  unsigned nblk;
  qggmrf_grid(S, R, C, &a.zchunk, &nblk);
  ktimer_begin(KC_QGGMRF, stream);
  hipLaunchKernelGGL((qggmrf_kernel<0, false>), dim3(nblk), dim3(256), 0, stream, a);
  TC_HIP(hipGetLastError());
  ktimer_end(KC_QGGMRF, stream, 3.0 * 4.0 * S * (double)R * C);
}

// ---------------------------------------------------------------------------
// free-boundary prior and its exact gradient (differentiable form)
// ---------------------------------------------------------------------------

__global__ void reduce_strided_kernel(const double* __restrict__ p, long count,
                                      int stride, double* out);

// rho(d) = u^2 / (c + u^(q-p)), u = |d| / sigma (q = 2)
__device__ __forceinline__ float rho_value(float d, const RhoParams& P) {
  const float u = fabsf(d) * P.inv_sigma;
  const float t = __builtin_amdgcn_exp2f(P.qmp * __builtin_amdgcn_logf(u));
  return u * u * __builtin_amdgcn_rcpf(MRF_C + t);
}

// MODE 0: g[v] += sum over in-bounds neighbours of w_d rho'(x_v - x_{v+d});
// MODE 1: per-block FP64 partial sums of the 13 forward pairs owned by v.
// One thread per voxel; neighbours come through the caches.
template <int MODE>
__global__ void __launch_bounds__(256)
qggmrf_free_kernel(const float* __restrict__ x, float* g, double* partials,
                   int S, int R, int C, RhoParams rp) {
  __shared__ double red[16];
  const long nvox = (long)S * R * C;
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (v < nvox) {
    const int z = (int)(v / ((long)R * C));
    const int rem = (int)(v - (long)z * R * C);
    const int r = rem / C, c = rem - (rem / C) * C;
    const float xv = x[v];
    float sum = 0.f;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int k = (dz != 0) + (dy != 0) + (dx != 0);
          if (k == 0) continue;
          // value: each unordered pair once, from its lexicographically
          // smaller end (forward offsets)
          const bool fwd = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
          if (MODE == 1 && !fwd) continue;
          const int zz = z + dz, rr = r + dy, cc = c + dx;
          if (zz < 0 || zz >= S || rr < 0 || rr >= R || cc < 0 || cc >= C)
            continue;
          const float w = k == 1 ? W_FACE : (k == 2 ? W_EDGE : W_CORNER);
          const float d = xv - x[((long)zz * R + rr) * C + cc];
          if (MODE == 0) {
            const float ws2 = w * rp.inv_sigma2;
            sum += wdrho(d, RhoW{ws2 * rp.qmp * MRF_C, ws2 * (2.0f - rp.qmp)},
                           rp);
          } else {
            acc += (double)(w * rho_value(d, rp));
          }
        }
    if (MODE == 0) g[v] += sum;
  }
  if (MODE == 1) {
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
  }
}

void add_qggmrf_grad_free(const float* x, float* g, int S, int R, int C,
                          float sigma, float p, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1, "qGGMRF: empty volume");
  const long nvox = (long)S * R * C;
  hipLaunchKernelGGL(qggmrf_free_kernel<0>, dim3(cdiv(nvox, 256)), dim3(256),
                     0, stream, x, g, nullptr, S, R, C, make_rho(sigma, p));
  TC_HIP(hipGetLastError());
}

void qggmrf_value_free_dev(const float* x, int S, int R, int C, float sigma,
                           float p, double* d_out, hipStream_t stream) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1, "qGGMRF: empty volume");
  const long nvox = (long)S * R * C;
  const unsigned nblk = (unsigned)cdiv(nvox, 256);
  DevBuf parts(sizeof(double) * nblk);
  hipLaunchKernelGGL(qggmrf_free_kernel<1>, dim3(nblk), dim3(256), 0, stream,
                     x, nullptr, parts.as<double>(), S, R, C,
                     make_rho(sigma, p));
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_strided_kernel, dim3(1), dim3(1024), 0, stream,
                     parts.as<double>(), (long)nblk, 1, d_out);
  TC_HIP(hipGetLastError());
}

double qggmrf_value_free(const float* x, int S, int R, int C, float sigma,
                         float p, hipStream_t stream) {
  DevBuf out(sizeof(double));
  qggmrf_value_free_dev(x, S, R, C, sigma, p, out.as<double>(), stream);
  double v = 0.0;
  TC_HIP(hipMemcpyAsync(&v, out.as<void>(), sizeof(double),
                        hipMemcpyDeviceToHost, stream));
  TC_HIP(hipStreamSynchronize(stream));
  return v;
}

float tv_hessian_zero(float sigma) {
  // accumulated in float like src/gpu/tv_hessian.cu:41-46
  const float d2 = 2.0f / (sigma * sigma * MRF_C);
  float acc = 0.f;
  for (int dz = 0; dz < 3; ++dz)
    for (int dy = 0; dy < 3; ++dy)
      for (int dx = 0; dx < 3; ++dx) {
        int k = (dz != 1) + (dy != 1) + (dx != 1);
        float w = k == 0 ? 0.f : (k == 1 ? W_FACE : (k == 2 ? W_EDGE : W_CORNER));
        acc += w * d2;
      }
  return acc;
}

// ---------------------------------------------------------------------------
// commit an accepted step:  xp <- y - s g,  Tx <- Ty - s Tg  (then swap)
// ---------------------------------------------------------------------------

// cand = y - s g and T cand = T y - s T g with y = sol + beta (sol - xp);
// dst / Tdst may alias sol / Tsol (element-wise in place), xp may alias sol.
__global__ void commit_kernel(const float* sol, const float* xp,
                              const float* Tsol, const float* Tx,
                              const float* __restrict__ g,
                              const float* __restrict__ Tg, float* dst,
                              float* Tdst, long n4, long n, float beta,
                              float s) {
  const long stride = (long)gridDim.x * blockDim.x;
  const float4* sol4 = reinterpret_cast<const float4*>(sol);
  const float4* xp4 = reinterpret_cast<const float4*>(xp);
  const float4* Tsol4 = reinterpret_cast<const float4*>(Tsol);
  const float4* Tx4 = reinterpret_cast<const float4*>(Tx);
  const float4* g4 = reinterpret_cast<const float4*>(g);
  const float4* Tg4 = reinterpret_cast<const float4*>(Tg);
  float4* d4 = reinterpret_cast<float4*>(dst);
  float4* Td4 = reinterpret_cast<float4*>(Tdst);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += stride) {
    const float4 so = sol4[i], x = xp4[i], gg = g4[i];
    const float4 ts = Tsol4[i], tx = Tx4[i], tg = Tg4[i];
    float4 o, ot;
    o.x = fmaf(-s, gg.x, fmaf(beta, so.x - x.x, so.x));
    o.y = fmaf(-s, gg.y, fmaf(beta, so.y - x.y, so.y));
    o.z = fmaf(-s, gg.z, fmaf(beta, so.z - x.z, so.z));
    o.w = fmaf(-s, gg.w, fmaf(beta, so.w - x.w, so.w));
    ot.x = fmaf(-s, tg.x, fmaf(beta, ts.x - tx.x, ts.x));
    ot.y = fmaf(-s, tg.y, fmaf(beta, ts.y - tx.y, ts.y));
    ot.z = fmaf(-s, tg.z, fmaf(beta, ts.z - tx.z, ts.z));
    ot.w = fmaf(-s, tg.w, fmaf(beta, ts.w - tx.w, ts.w));
    d4[i] = o;
    Td4[i] = ot;
  }
  for (long i = 4 * n4 + (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const float so = sol[i], x = xp[i], ts = Tsol[i], tx = Tx[i];
    dst[i] = fmaf(-s, g[i], fmaf(beta, so - x, so));
    Tdst[i] = fmaf(-s, Tg[i], fmaf(beta, ts - tx, ts));
  }
}

// out[k] = sum_i p[i*stride + k] for k < nvals (one workgroup per value)
__global__ void reduce_strided_kernel(const double* __restrict__ p, long count,
                                      int stride, double* out) {
  __shared__ double red[16];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (long i = threadIdx.x; i < count; i += blockDim.x)
    acc += p[i * stride + k];
  double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[k] = tot;
}

// The line search's sums in one launch: out[k] = sum_i q[i*nq + k] for
// k < nq (the gradient kernel's per-workgroup dots), out[nq] = sum_i t[i]
// (the Toeplitz column pass's g.Tg partials).  Same order per value as
// reduce_strided_kernel.
__global__ void reduce_dots_kernel(const double* __restrict__ q, long qcount,
                                   int nq, const double* __restrict__ t,
                                   long tcount, double* out) {
  __shared__ double red[16];
  const int k = blockIdx.x;
  const double* p = k < nq ? q + k : t;
  const long count = k < nq ? qcount : tcount;
  const int stride = k < nq ? nq : 1;
  // eight independent partial sums keep eight loads in flight per thread
  // (one block sums tens of thousands of Toeplitz partials); fixed order.
  // 22.7 -> 12.5 us per attempt at 32 x 1454^2.
  const long bd = blockDim.x;
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  long i = threadIdx.x;
  for (; i + 7 * bd < count; i += 8 * bd) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += p[(i + u * bd) * stride];
  }
  for (; i < count; i += bd) a[0] += p[i * stride];
  const double acc = ((a[0] + a[1]) + (a[2] + a[3])) +
                     ((a[4] + a[5]) + (a[6] + a[7]));
  double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[k] = tot;
}

// ---------------------------------------------------------------------------
// per-slice scales of the FP16 gradient (mixed precision)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float block_max256(float m, float* red) {
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// power of two mapping m into (2^11, 2^12]; `keep` when m is 0 or not finite
__device__ __forceinline__ float half_scale_for(float m, float keep) {
  if (!(m > 0.f && m <= 3.0e38f)) return keep;
  int ex = 0;
  (void)frexpf(m, &ex);
  return ldexpf(1.f, 12 - max(ex, -100));
}

// max |x| of each slice into mx[s] (bit patterns of non-negative floats
// order like the floats, so an integer atomicMax combines the workgroups;
// mx starts at 0): many workgroups per slice, where one per slice left the
// GPU mostly idle (13.5 ms for 64 x 2908^2)
__global__ void __launch_bounds__(256)
slice_absmax_kernel(const float* __restrict__ x, long plane, unsigned* mx) {
  __shared__ float red[4];
  const float* p = x + (long)blockIdx.y * plane;
  float m = 0.f;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < plane;
       i += (long)gridDim.x * 256)
    m = fmaxf(m, fabsf(p[i]));
  m = block_max256(m, red);
  if (threadIdx.x == 0) atomicMax(&mx[blockIdx.y], __float_as_uint(m));
}

__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// gsc[s] from max |x[s]|
__global__ void gscale_from_max_kernel(const unsigned* __restrict__ mx, int S,
                                       float* gsc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) gsc[s] = half_scale_for(__uint_as_float(mx[s]), 1.f);
}

static void slice_scales(const float* x, long plane, int S, float* gsc,
                         hipStream_t stream) {
  DevBuf mx(sizeof(unsigned) * (size_t)S);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(cdiv(S, 256)), dim3(256), 0,
                     stream, mx.as<unsigned>(), S, 0u);
  TC_HIP(hipGetLastError());
  const unsigned nx =
      (unsigned)std::max<long>(1, std::min<long>(256, cdiv(plane, 256L * 16)));
  hipLaunchKernelGGL(slice_absmax_kernel, dim3(nx, S), dim3(256), 0, stream,
                     x, plane, mx.as<unsigned>());
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(gscale_from_max_kernel, dim3(cdiv(S, 256)), dim3(256), 0,
                     stream, mx.as<unsigned>(), S, gsc);
  TC_HIP(hipGetLastError());
}

// next[s] from the largest |g| of slice s this attempt (the stencil's
// per-wave maxima); counts slices whose stored values saturated
__global__ void __launch_bounds__(256)
gscale_update_kernel(const float* __restrict__ gmaxp, int ntiles, int zchunk,
                     const float* __restrict__ cur, float* next,
                     int* saturated) {
  __shared__ float red[4];
  const int s = blockIdx.x;
  const int zc = s / zchunk, off = s - zc * zchunk;
  const long nw = (long)ntiles * QWAVES;
  const long base = (long)zc * nw;
  float m = 0.f;
  for (long i = threadIdx.x; i < nw; i += 256)
    m = fmaxf(m, gmaxp[(base + i) * zchunk + off]);
  m = block_max256(m, red);
  if (threadIdx.x == 0) {
    const float c = cur[s];
    next[s] = half_scale_for(m, c);
    if (!(m * c <= 65504.f)) atomicAdd(saturated, 1);
  }
}

size_t half_pair_bytes(int S, int R, int C) {
  return sizeof(unsigned) * (size_t)S * (size_t)((R + 1) / 2) * (size_t)C;
}

__global__ void to_half_slices_kernel(const float* __restrict__ x,
                                      unsigned* out, int R, int C, long n,
                                      const float* __restrict__ gsc) {
  // one word (two rows of a column) per thread
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int np = (R + 1) / 2;
  const long sq = i / C;
  const int c = (int)(i - sq * C);
  const int s = (int)(sq / np);
  const int q = (int)(sq - (long)s * np);
  const float g = gsc[s];
  const float* xs = x + (long)s * R * C;
  auto h = [&](int r) {
    return r < R ? __float2half_rn(fminf(fmaxf(xs[(long)r * C + c] * g,
                                               -65504.f), 65504.f))
                 : __float2half_rn(0.f);
  };
  out[i] = pack_h2(h(2 * q), h(2 * q + 1));
}
void to_half_slices(const float* x, void* x16, float* gsc, int S, int R,
                    int C, hipStream_t stream) {
  slice_scales(x, (long)R * C, S, gsc, stream);
  const long n = (long)S * ((R + 1) / 2) * C;
  hipLaunchKernelGGL(to_half_slices_kernel, dim3(cdiv(n, 256)), dim3(256), 0,
                     stream, x, static_cast<unsigned*>(x16), R, C, n, gsc);
  TC_HIP(hipGetLastError());
}

// Optimizer::run2 (src/optimize.h:96-154), restated on device-resident state.
// Per attempt the reference evaluates gradient_(y), error_(sol), error_(y)
// (and error_(sol) again after acceptance): four Toeplitz applications.  By
// linearity of T only T g is new, because
//   T y    = T sol + beta (T sol - T x)
//   T cand = T y - s T g
// and err^2(v) = ||y_data||^2 + v . T v - 2 v . b expands into dot products
// of y, g, T y, T g, b.  Semantics kept from the reference: `sol` takes the
// candidate on every attempt (accepted or not), `x` takes it only on
// acceptance, so after an acceptance x == sol and the next first attempt has
// y = sol.
// The stream the MBIR loop runs on.  HIP graphs need a stream that can be
// captured (not the null stream), so with graphs on the loop runs on a
// private non-blocking stream (default priority) ordered after the caller's
// work; the destructor drains it before the buffers declared earlier are
// freed.
struct LoopStream {
  hipStream_t s = nullptr;
  bool own = false;
  LoopStream(hipStream_t caller, bool private_stream) : s(caller) {
    if (!private_stream) return;
    hipEvent_t ev;
    TC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) {
      own = true;
      e = hipEventRecord(ev, caller);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, ev, 0);
    }
    (void)hipEventDestroy(ev);
    if (e != hipSuccess) {
      if (own) (void)hipStreamDestroy(s);
      throw std::runtime_error(std::string("HIP error: ") +
                               hipGetErrorString(e));
    }
  }
  ~LoopStream() {
    if (!own) return;
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  LoopStream(const LoopStream&) = delete;
  LoopStream& operator=(const LoopStream&) = delete;
};

// Captures the work `enqueue` puts on `stream` into an executable graph;
// nullptr (and no error left pending) if any step of the capture fails.
template <class F>
static hipGraphExec_t capture_attempt(hipStream_t stream, F&& enqueue) {
  if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) !=
      hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  hipGraph_t graph = nullptr;
  bool ok = true;
  try {
    (void)enqueue();
  } catch (...) {
    ok = false;
  }
  if (hipStreamEndCapture(stream, &graph) != hipSuccess) ok = false;
  hipGraphExec_t exec = nullptr;
  if (ok && graph &&
      hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    exec = nullptr;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  return exec;
}

static std::atomic<long> g_graph_replays{0};
long graph_replays() { return g_graph_replays.load(); }

// Instantiated graphs of one steady-state attempt, keyed by the buffer
// roles and the step they were captured with (the step is a kernel
// argument of the capture; a different step is a miss and runs eagerly).
struct AttemptGraphs {
  struct Entry {
    const float* sol;
    const float* Tsol;
    const float* gsc;
    float step;
    bool fused;
    hipGraphExec_t exec;
  };
  std::vector<Entry> e;
  Entry* find(const float* sol, const float* Tsol, const float* gsc,
              float step) {
    for (auto& x : e)
      if (x.sol == sol && x.Tsol == Tsol && x.gsc == gsc && x.step == step)
        return &x;
    return nullptr;
  }
  ~AttemptGraphs() {
    for (auto& x : e)
      if (x.exec) (void)hipGraphExecDestroy(x.exec);
  }
};

Run2Stats run2_device(Psf& psf, const float* x0, const float* b, float* out,
                      int S, const Run2Options& opt, hipStream_t caller) {
  const int n = psf.n();
  TC_REQUIRE(S >= 1, "run2: need at least one slice");
  TC_REQUIRE(opt.num_iters >= 0, "run2: num_iters must be >= 0");
  const long N = (long)S * n * n;
  const size_t vb = sizeof(float) * (size_t)N;
  // buffers from the thread's arena (the loop ends with a synchronisation)
  ArenaScope arena;
  DevBuf bA(vb), bB(vb), bTA(vb), bTB(vb), bg(vb), bTg(vb);
  float* A = bA.as<float>();
  float* B = bB.as<float>();
  float* TA = bTA.as<float>();
  float* TB = bTB.as<float>();
  float* g = bg.as<float>();
  float* Tg = bTg.as<float>();

  QArgs qa{};
  qa.b = b;
  qa.g = g;
  qa.S = S;
  qa.R = n;
  qa.C = n;
  set_rho(qa, opt.sigma, opt.p);
  unsigned qblk;
  qggmrf_grid(S, n, n, &qa.zchunk, &qblk);
  const int ntp = toeplitz_num_partials(psf, S);
  DevBuf parts(sizeof(double) * ((size_t)qblk * NDOT + ntp));
  double* qparts = parts.as<double>();
  double* tparts = qparts + (size_t)qblk * NDOT;
  qa.partials = qparts;
  // mixed precision: FP16 gradient, per-slice scales (this attempt's and
  // the next one's), per-wave maxima, saturation counter
  DevBuf bg16, bgsc, bgmax, bsat;
  DevBuf dots_dev(sizeof(double) * 8);
  // pinned landing area of the per-attempt dots (the arena's, kept with it)
  double* dots_h = arena_pinned_dots();
  // TOMOCAM_HIP_FUSED_COMMIT=0 selects the separate commit kernel (same
  // arithmetic, one more pass over HBM)
  const char* fenv = std::getenv("TOMOCAM_HIP_FUSED_COMMIT");
  const bool allow_fused = !(fenv && fenv[0] == '0');
  // With TOMOCAM_HIP_GRAPHS=1, steady-state attempts (y = sol, step s0)
  // replay a HIP graph per buffer role.  Not with kernel timing (its events
  // are per launch), the separate commit kernel or a second Toeplitz
  // stream.  Off by default: per call it adds a stream and up to eight
  // captures and instantiations, and per attempt a graph launch measured
  // no faster than the 6-15 launches it replaces (4 x 362^2, 50
  // iterations: 7.8 ms against 7.2 ms eager; 32 x 1454^2: equal).
  const char* genv = std::getenv("TOMOCAM_HIP_GRAPHS");
  bool use_graphs = opt.allow_graphs && genv && genv[0] == '1' &&
                    !ktimer_active() && psf.has_fused() &&
                    (allow_fused || opt.mixed) &&
                    psf.single_stream(S, false);  // FP32 intermediates
  AttemptGraphs graphs;
  LoopStream loop_stream(caller, use_graphs);
  const hipStream_t stream = loop_stream.s;
  float *gsc = nullptr, *gsc_next = nullptr;
  if (opt.mixed) {
    TC_REQUIRE(psf.has_fused(),
               "mixed precision needs a compile-time FFT plan; FFT length " +
                   std::to_string(psf.M()) + " has none");
    bg16.alloc(half_pair_bytes(S, n, n));
    bgsc.alloc(sizeof(float) * 2 * (size_t)S);
    bgmax.alloc(sizeof(float) * (size_t)qblk * QWAVES * qa.zchunk);
    bsat.alloc(sizeof(int));
    TC_HIP(hipMemsetAsync(bsat.as<void>(), 0, sizeof(int), stream));
    gsc = bgsc.as<float>();
    gsc_next = gsc + S;
    qa.g16 = bg16.as<unsigned>();
    qa.gmaxp = bgmax.as<float>();
    // the first gradient is about -b (x0 is small): scales from max |b|
    slice_scales(b, (long)n * n, S, gsc, stream);
  }
  const int ntiles_q = (int)(qblk / cdiv(S, qa.zchunk));

  // sol = x = x0, T sol = T x = T x0 (aliased until the first rejection)
  TC_HIP(hipMemcpyAsync(A, x0, vb, hipMemcpyDeviceToDevice, stream));
  psf.apply(A, nullptr, TA, S, 0, nullptr, stream);
  float *sol = A, *xp = A, *Tsol = TA, *Tx = TA;

  Run2Stats st;
  st.toeplitz_applies = 1;
  const unsigned cblk = 2048;
  double t = 1.0, tnew = 1.0;
  const double step0 = opt.step;
  double step = step0;
  const double yd2 = opt.sino_norm;
  double xerr = (double)N;
  for (int iter = 0; iter < opt.num_iters; ++iter) {
    double fx = 0.0;
    int bt_this = 0;
    for (;;) {
      const double beta = tnew * (1.0 / t - 1.0);
      tnew = 0.5 * (std::sqrt(t * t * t * t + 4.0 * t * t) - t * t);
      const bool has_x = sol != xp;
      qa.sol = sol;
      qa.xprev = has_x ? xp : sol;
      qa.Tsol = Tsol;
      qa.Tx = has_x ? Tx : Tsol;
      qa.beta = has_x ? (float)beta : 0.0f;
      qa.gsc = gsc;
      // The step is committed inside the Toeplitz passes where the FFT size
      // allows: cand = y - s g and T cand = T y - s T g are written before
      // the accept test, which then only decides the buffers' roles.  In
      // place unless x aliases sol (then sol must survive a rejection).
      float* fdst = has_x ? sol : (sol == A ? B : A);
      float* fTdst = has_x ? Tsol : (Tsol == TA ? TB : TA);
      const Psf::StepCommit sc{sol, has_x ? xp : nullptr, Tsol,
                               has_x ? Tx : nullptr, fdst, fTdst,
                               has_x ? (float)beta : 0.0f, (float)step};
      // one attempt's work: stencil (+ FP16 scales), Toeplitz passes with
      // the commit, dot reduction, dots to the host; returns whether the
      // commit was fused
      auto enqueue = [&]() -> bool {
        ktimer_begin(KC_MBIR_GRADIENT, stream);
        if (opt.mixed) {
          if (has_x)
            hipLaunchKernelGGL((qggmrf_kernel<1, true, true>), dim3(qblk),
                               dim3(256), 0, stream, qa);
          else
            hipLaunchKernelGGL((qggmrf_kernel<1, false, true>), dim3(qblk),
                               dim3(256), 0, stream, qa);
        } else if (has_x) {
          hipLaunchKernelGGL((qggmrf_kernel<1, true>), dim3(qblk), dim3(256),
                             0, stream, qa);
        } else {
          hipLaunchKernelGGL((qggmrf_kernel<1, false>), dim3(qblk), dim3(256),
                             0, stream, qa);
        }
        TC_HIP(hipGetLastError());
        // reads sol, T sol, b (and x, T x unless aliased), writes g
        ktimer_end(KC_MBIR_GRADIENT, stream,
                   (has_x ? 5.0 : 3.0) * vb + (opt.mixed ? 0.5 : 1.0) * vb);
        if (opt.mixed) {
          hipLaunchKernelGGL(gscale_update_kernel, dim3(S), dim3(256), 0,
                             stream, qa.gmaxp, ntiles_q, qa.zchunk, gsc,
                             gsc_next, bsat.as<int>());
          TC_HIP(hipGetLastError());
        }
        const bool f =
            opt.mixed
                ? psf.apply_commit_half(qa.g16, gsc, sc, S, tparts, stream,
                                        /*wt16=*/false)
                : allow_fused && psf.apply_commit(g, sc, S, tparts, stream);
        if (!f) psf.apply(g, nullptr, Tg, S, 0, tparts, stream);
        hipLaunchKernelGGL(reduce_dots_kernel, dim3(NDOT + 1), dim3(1024), 0,
                           stream, qparts, (long)qblk, NDOT, tparts,
                           (long)ntp, dots_dev.as<double>());
        TC_HIP(hipGetLastError());
        TC_HIP(hipMemcpyAsync(dots_h, dots_dev.as<double>(),
                              sizeof(double) * 8, hipMemcpyDeviceToHost,
                              stream));
        return f;
      };
      // Steady state (y = sol, step s0): the first attempt with given buffer
      // roles runs eagerly (first-use allocations happen there), the second
      // captures a graph, later ones replay it.  A capture that fails turns
      // graphs off for the rest of the call.
      bool fused;
      AttemptGraphs::Entry* ge =
          use_graphs && !has_x ? graphs.find(sol, Tsol, gsc, (float)step)
                               : nullptr;
      if (ge && !ge->exec) {
        ge->exec = capture_attempt(stream, enqueue);
        if (!ge->exec) use_graphs = false;
      }
      if (ge && ge->exec) {
        TC_HIP(hipGraphLaunch(ge->exec, stream));
        g_graph_replays.fetch_add(1);
        fused = ge->fused;
      } else {
        fused = enqueue();
        if (use_graphs && !has_x && !ge)
          graphs.e.push_back({sol, Tsol, gsc, (float)step, fused, nullptr});
      }
      if (opt.mixed) std::swap(gsc, gsc_next);
      ++st.toeplitz_applies;
      TC_HIP(hipStreamSynchronize(stream));
      ++st.attempts;
      const double yTy = dots_h[0], yb = dots_h[1], gR = dots_h[2];
      const double gg = dots_h[3], dd = dots_h[4], dg = dots_h[5];
      const double gTg = dots_h[NDOT];
      const double s = step;
      // err^2(y) = ||y_data||^2 + y.Ty - 2 y.b; err^2(cand) - err^2(y)
      // = s (s g.Tg - 2 g.(Ty - b)) is formed directly (no cancellation)
      const double fy2 = yd2 + yTy - 2.0 * yb;
      const double dlt = s * (s * gTg - 2.0 * gR);
      const double fy = std::sqrt(fy2);
      fx = std::sqrt(fy2 + dlt);
      const double gy = 0.5 * s * gg;
      // fx > fy + gy  <=>  (fx - fy) > gy, with fx - fy = dlt / (fx + fy);
      // NaN compares false (accept) as in the reference
      const bool reject = (dlt / (fx + fy)) > gy;
      float* dst = fdst;
      float* Tdst = fTdst;
      if (!fused) {
        // cand = y - s g, T cand = T y - s T g; in place unless x aliases
        // sol and must survive a rejection
        dst = (reject && !has_x) ? (sol == A ? B : A) : sol;
        Tdst = (reject && !has_x) ? (Tsol == TA ? TB : TA) : Tsol;
        const float cb = has_x ? (float)beta : 0.0f;
        ktimer_begin(KC_MBIR_COMMIT, stream);
        hipLaunchKernelGGL(commit_kernel, dim3(cblk), dim3(256), 0, stream,
                           sol, qa.xprev, Tsol, qa.Tx, g, Tg, dst, Tdst, N / 4,
                           N, cb, (float)s);
        TC_HIP(hipGetLastError());
        // reads sol, T sol, g, T g (and x, T x unless aliased), writes 2
        ktimer_end(KC_MBIR_COMMIT, stream, (has_x ? 8.0 : 6.0) * vb);
      }
      sol = dst;
      Tsol = Tdst;
      if (reject) {
        step *= 0.9;
        ++st.backtracks;
        ++bt_this;
        continue;
      }
      const double opb = 1.0 + (has_x ? beta : 0.0);
      xerr = opb * opb * dd - 2.0 * opb * s * dg + s * s * gg;
      xp = sol;
      Tx = Tsol;
      const double s_acc = s;
      step = step0;
      t = tnew;
      ++st.iterations;
      if (opt.refresh_every > 0 && st.iterations % opt.refresh_every == 0 &&
          iter + 1 < opt.num_iters) {
        // T sol from scratch (x == sol here), bounding FP32 drift of the
        // recurrence
        float* fresh = (Tsol == TA) ? TB : TA;
        psf.apply(sol, nullptr, fresh, S, 0, nullptr, stream);
        Tsol = Tx = fresh;
        ++st.toeplitz_applies;
      }
      if (opt.on_iter) opt.on_iter(iter, fx, std::sqrt(xerr), s_acc, bt_this);
      break;
    }
    st.last_error = fx;
    st.last_xerr = std::sqrt(xerr);
    if (opt.verbose)
      std::fprintf(stdout, "iter: %4d, error: %5.4e, x-err: %5.4e\n", iter, fx,
                   std::sqrt(xerr));
  }
  TC_HIP(hipMemcpyAsync(out, sol, vb, hipMemcpyDeviceToDevice, stream));
  if (opt.mixed)
    TC_HIP(hipMemcpyAsync(&st.saturated, bsat.as<int>(), sizeof(int),
                          hipMemcpyDeviceToHost, stream));
  TC_HIP(hipStreamSynchronize(stream));
  return st;
}

}  // namespace tc
