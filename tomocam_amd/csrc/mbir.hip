// qGGMRF stencil, fused MBIR gradient/commit kernels and the run2 driver.
#include <cmath>
#include <cstdio>
#include <utility>

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

struct RhoParams {
  float inv_sigma;
  float inv_sigma2;
  float qmp;  // q - p
};

__device__ __forceinline__ float drho(float d, const RhoParams& P) {
  const float a = fabsf(d) * P.inv_sigma;
  // (|d|/sigma)^(q-p) = 2^((q-p) log2(|d|/sigma)); log2(0) = -inf -> t = 0
  const float t = __builtin_amdgcn_exp2f(P.qmp * __builtin_amdgcn_logf(a));
  const float r = __builtin_amdgcn_rcpf(MRF_C + t);
  return d * P.inv_sigma2 * r * (2.0f - P.qmp * t * r);
}

struct QArgs {
  const float* sol;    // y = sol (+ beta (sol - xprev) in MBIR mode)
  const float* xprev;
  const float* Tsol;
  const float* Tx;
  const float* b;
  float* g;
  double* partials;    // NDOT per workgroup (MBIR mode)
  int S, R, C;
  int zchunk;
  float beta;
  RhoParams rp;
};

constexpr int QTX = 64, QTY = 8, QPW = QTX + 2, QPH = QTY + 2;
// line-search dot products per workgroup (MBIR mode):
//   y.Ty, y.b, g.(Ty - b), g.g, d.d, d.g   (d = sol - x)
constexpr int NDOT = 6;

// MODE 0: g += R'(x).  MODE 1 (MBIR): y = sol + beta (sol - xprev),
// g = Tsol + beta (Tsol - Tx) - b + R'(y), plus the line-search dots.
template <int MODE>
__global__ void __launch_bounds__(256) qggmrf_kernel(QArgs a) {
  __shared__ float pl[3][QPH * QPW];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int ntx = (a.C + QTX - 1) / QTX;
  const int nty = (a.R + QTY - 1) / QTY;
  const int ntiles = ntx * nty;
  const int bl = xcd_remap(blockIdx.x, gridDim.x);
  const int zc = bl / ntiles;
  const int tile = bl - zc * ntiles;
  const int ty = tile / ntx;
  const int tx = tile - ty * ntx;
  const int r0 = ty * QTY, c0 = tx * QTX;
  const int z0 = zc * a.zchunk;
  const int z1 = min(a.S, z0 + a.zchunk);
  const int lx = tid & 63;
  const int ly = tid >> 6;
  const long plane = (long)a.R * a.C;

  auto load = [&](int slot, int z) {
    const int zz = min(max(z, 0), a.S - 1);
    const long zo = (long)zz * plane;
    for (int t = tid; t < QPH * QPW; t += 256) {
      int rr = t / QPW;
      int cc = t - rr * QPW;
      int r = min(max(r0 + rr - 1, 0), a.R - 1);
      int c = min(max(c0 + cc - 1, 0), a.C - 1);
      long o = zo + (long)r * a.C + c;
      float v = a.sol[o];
      if (MODE == 1) v = fmaf(a.beta, v - a.xprev[o], v);
      pl[slot][t] = v;
    }
  };

  int sm1 = 0, s0 = 1, sp1 = 2;
  load(sm1, z0 - 1);
  load(s0, z0);
  load(sp1, z0 + 1);
  __syncthreads();

  double acc[NDOT] = {0, 0, 0, 0, 0, 0};
  for (int z = z0; z < z1; ++z) {
    const float* P0 = pl[sm1];
    const float* P1 = pl[s0];
    const float* P2 = pl[sp1];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int lr = ly + 4 * half;
      const int r = r0 + lr, c = c0 + lx;
      if (r < a.R && c < a.C) {
        const int ctr = (lr + 1) * QPW + (lx + 1);
        const float v = P1[ctr];
        float sum = 0.f;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) {
          const float* Pz = dz == 0 ? P0 : (dz == 1 ? P1 : P2);
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
              const int k = (dz != 1) + (dy != 1) + (dx != 1);
              if (k == 0) continue;
              const float w = k == 1 ? W_FACE : (k == 2 ? W_EDGE : W_CORNER);
              sum += w * drho(v - Pz[ctr + (dy - 1) * QPW + (dx - 1)], a.rp);
            }
          }
        }
        const long o = (long)z * plane + (long)r * a.C + c;
        if (MODE == 0) {
          a.g[o] += sum;
        } else {
          const float so = a.sol[o], xp = a.xprev[o];
          const float ts = a.Tsol[o], tq = a.Tx[o], bb = a.b[o];
          const float tyv = fmaf(a.beta, ts - tq, ts);
          const float gv = (tyv - bb) + sum;
          a.g[o] = gv;
          const float d = so - xp;
          acc[0] += (double)v * tyv;
          acc[1] += (double)v * bb;
          acc[2] += (double)gv * (tyv - bb);
          acc[3] += (double)gv * gv;
          acc[4] += (double)d * d;
          acc[5] += (double)d * gv;
        }
      }
    }
    __syncthreads();
    const int t = sm1;
    sm1 = s0;
    s0 = sp1;
    sp1 = t;
    if (z + 1 < z1) load(sp1, z + 2);
    __syncthreads();
  }
  if (MODE == 1) {
#pragma unroll
    for (int k = 0; k < NDOT; ++k) {
      double tot = block_sum(acc[k], red);
      if (tid == 0) a.partials[(long)bl * NDOT + k] = tot;
    }
  }
}

static RhoParams make_rho(float sigma, float p) {
  TC_REQUIRE(sigma > 0.f && std::isfinite(sigma), "qGGMRF: sigma must be > 0");
  RhoParams r;
  r.inv_sigma = 1.0f / sigma;
  r.inv_sigma2 = 1.0f / (sigma * sigma);
  r.qmp = 2.0f - p;
  return r;
}

static void qggmrf_grid(int S, int R, int C, int* zchunk, unsigned* nblk) {
  const int ntiles = (int)(cdiv(C, QTX) * cdiv(R, QTY));
  int zc = S;
  // keep at least ~2048 workgroups in flight
  while (zc > 4 && (long)ntiles * cdiv(S, zc) < 2048) zc = (zc + 1) / 2;
  *zchunk = zc;
  *nblk = (unsigned)(ntiles * cdiv(S, zc));
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
  a.rp = make_rho(sigma, p);
  unsigned nblk;
  qggmrf_grid(S, R, C, &a.zchunk, &nblk);
  ktimer_begin(KC_QGGMRF, stream);
  hipLaunchKernelGGL(qggmrf_kernel<0>, dim3(nblk), dim3(256), 0, stream, a);
  TC_HIP(hipGetLastError());
  ktimer_end(KC_QGGMRF, stream, 3.0 * 4.0 * S * (double)R * C);
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
Run2Stats run2_device(Psf& psf, const float* x0, const float* b, float* out,
                      int S, const Run2Options& opt, hipStream_t stream) {
  const int n = psf.n();
  TC_REQUIRE(S >= 1, "run2: need at least one slice");
  TC_REQUIRE(opt.num_iters >= 0, "run2: num_iters must be >= 0");
  const long N = (long)S * n * n;
  const size_t vb = sizeof(float) * (size_t)N;
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
  qa.rp = make_rho(opt.sigma, opt.p);
  unsigned qblk;
  qggmrf_grid(S, n, n, &qa.zchunk, &qblk);
  const int ntp = toeplitz_num_partials(psf, S);
  DevBuf parts(sizeof(double) * ((size_t)qblk * NDOT + ntp));
  double* qparts = parts.as<double>();
  double* tparts = qparts + (size_t)qblk * NDOT;
  qa.partials = qparts;
  DevBuf dots_dev(sizeof(double) * 8);
  double* dots_h = nullptr;
  TC_HIP(hipHostMalloc((void**)&dots_h, sizeof(double) * 8, hipHostMallocDefault));
  struct HostFree {
    double* p;
    ~HostFree() { if (p) (void)hipHostFree(p); }
  } hf{dots_h};

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
      ktimer_begin(KC_MBIR_GRADIENT, stream);
      hipLaunchKernelGGL(qggmrf_kernel<1>, dim3(qblk), dim3(256), 0, stream, qa);
      TC_HIP(hipGetLastError());
      // reads sol, T sol, b (and x, T x unless aliased), writes g
      ktimer_end(KC_MBIR_GRADIENT, stream, (has_x ? 6.0 : 4.0) * vb);
      psf.apply(g, nullptr, Tg, S, 0, tparts, stream);
      ++st.toeplitz_applies;
      hipLaunchKernelGGL(reduce_strided_kernel, dim3(NDOT), dim3(1024), 0,
                         stream, qparts, (long)qblk, NDOT,
                         dots_dev.as<double>());
      hipLaunchKernelGGL(reduce_strided_kernel, dim3(1), dim3(1024), 0, stream,
                         tparts, (long)ntp, 1, dots_dev.as<double>() + NDOT);
      TC_HIP(hipGetLastError());
      TC_HIP(hipMemcpyAsync(dots_h, dots_dev.as<double>(), sizeof(double) * 8,
                            hipMemcpyDeviceToHost, stream));
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
      // cand = y - s g, T cand = T y - s T g; in place unless x aliases sol
      // and must survive a rejection
      float* dst = (reject && !has_x) ? (sol == A ? B : A) : sol;
      float* Tdst = (reject && !has_x) ? (Tsol == TA ? TB : TA) : Tsol;
      const float cb = has_x ? (float)beta : 0.0f;
      ktimer_begin(KC_MBIR_COMMIT, stream);
      hipLaunchKernelGGL(commit_kernel, dim3(cblk), dim3(256), 0, stream, sol,
                         qa.xprev, Tsol, qa.Tx, g, Tg, dst, Tdst, N / 4, N, cb,
                         (float)s);
      TC_HIP(hipGetLastError());
      // reads sol, T sol, g, T g (and x, T x unless aliased), writes 2
      ktimer_end(KC_MBIR_COMMIT, stream, (has_x ? 8.0 : 6.0) * vb);
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
  TC_HIP(hipStreamSynchronize(stream));
  return st;
}

}  // namespace tc
