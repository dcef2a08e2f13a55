// FP64 forms of the hot path (declarations and scope in fp64.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "driver.h"
#include "fp64.h"

namespace tc {
namespace f64 {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 zadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 zsub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// ---------------------------------------------------------------------------
// FFT: runtime-planned Stockham transform of lines held in LDS
// ---------------------------------------------------------------------------

constexpr int kMaxStages = 24;

struct Plan {
  int L;
  int nst;
  int radix[kMaxStages];
  const double2* tw;  // tw[m] = exp(-2 pi i m / L)
};

static std::vector<int> factor(int L) {
  TC_REQUIRE(L >= 1 && L <= kMaxFFT,
             "FP64 FFT length " + std::to_string(L) + " outside [1, " +
                 std::to_string(kMaxFFT) + "]");
  std::vector<int> r;
  int m = L;
  while (m % 4 == 0) {
    r.push_back(4);
    m /= 4;
  }
  for (int p : {2, 3, 5, 7})
    while (m % p == 0) {
      r.push_back(p);
      m /= p;
    }
  TC_REQUIRE(m == 1, "FP64 FFT length " + std::to_string(L) +
                         " is not a product of 2, 3, 5, 7");
  TC_REQUIRE((int)r.size() <= kMaxStages, "FP64 FFT length too large");
  return r;
}

static int smooth_at_least(int n) {
  for (int m = std::max(n, 1);; ++m) {
    int r = m;
    for (int p : {2, 3, 5, 7})
      while (r % p == 0) r /= p;
    if (r == 1) return m;
  }
}

class Length {
 public:
  explicit Length(int L) {
    ArenaPause no_arena;  // process-lifetime table, filled synchronously
    std::vector<int> r = factor(L);
    p_.L = L;
    p_.nst = (int)r.size();
    for (int i = 0; i < p_.nst; ++i) p_.radix[i] = r[i];
    std::vector<double2> h(L);
    for (int m = 0; m < L; ++m) {
      const double a = -2.0 * M_PI * (double)m / (double)L;
      h[m] = make_double2(std::cos(a), std::sin(a));
    }
    tw_.alloc(sizeof(double2) * L);
    TC_HIP(hipMemcpy(tw_.as<void>(), h.data(), sizeof(double2) * L,
                     hipMemcpyHostToDevice));
    p_.tw = tw_.as<double2>();
  }
  const Plan& plan() const { return p_; }

 private:
  Plan p_;
  DevBuf tw_;
};

// exp(-2 pi i m / N), m < N, for any N (the ray DFT's roots)
class Roots {
 public:
  explicit Roots(int N) {
    ArenaPause no_arena;  // process-lifetime table, filled synchronously
    std::vector<double2> h(N);
    for (int m = 0; m < N; ++m) {
      const double a = -2.0 * M_PI * (double)m / (double)N;
      h[m] = make_double2(std::cos(a), std::sin(a));
    }
    t_.alloc(sizeof(double2) * N);
    TC_HIP(hipMemcpy(t_.as<void>(), h.data(), sizeof(double2) * N,
                     hipMemcpyHostToDevice));
  }
  const double2* table() const { return t_.as<double2>(); }

 private:
  DevBuf t_;
};

static const double2* roots(int N) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::unique_ptr<Roots>> cache;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, N);
  auto it = cache.find(key);
  if (it == cache.end())
    it = cache.emplace(key, std::make_unique<Roots>(N)).first;
  return it->second->table();
}

static const Length& length(int L) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::unique_ptr<Length>> cache;
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, L);
  auto it = cache.find(key);
  if (it == cache.end())
    it = cache.emplace(key, std::make_unique<Length>(L)).first;
  return *it->second;
}

// radix-R DFT of a[0..R) in place; roots w^k = tw[k L / R] (DIR = -1) or
// their conjugates (DIR = +1)
template <int R, int DIR>
__device__ __forceinline__ void dft_small(double2* a,
                                          const double2* __restrict__ tw,
                                          int L) {
  if constexpr (R == 2) {
    const double2 t = a[1];
    a[1] = zsub(a[0], t);
    a[0] = zadd(a[0], t);
  } else if constexpr (R == 4) {
    const double2 t0 = zadd(a[0], a[2]), t1 = zsub(a[0], a[2]);
    const double2 t2 = zadd(a[1], a[3]), d = zsub(a[1], a[3]);
    // d * (DIR i)
    const double2 t3 = DIR < 0 ? make_double2(d.y, -d.x)
                               : make_double2(-d.y, d.x);
    a[0] = zadd(t0, t2);
    a[2] = zsub(t0, t2);
    a[1] = zadd(t1, t3);
    a[3] = zsub(t1, t3);
  } else {
    double2 out[R];
    const int step = L / R;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      double2 acc = a[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const double2 w = tw[((r * k) % R) * step];
        acc = zadd(acc, DIR < 0 ? zmul(a[r], w) : zmulc(a[r], w));
      }
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) a[k] = out[k];
  }
}

// One Stockham stage of radix R on the LDS line: each thread holds up to 4
// butterflies (read, twiddle, DFT), then writes them after a barrier.
template <int R, int DIR>
__device__ __forceinline__ void lds_stage(double2* buf, int L, int Ns,
                                          const double2* __restrict__ tw) {
  const int nb = L / R;
  const int nthr = blockDim.x;
  double2 v[4][R];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = threadIdx.x + q * nthr;
    if (j < nb) {
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double2 x = buf[j + r * nb];
        if (r > 0 && Ns > 1) {
          const double2 w = tw[(long)r * k * (L / (Ns * R))];
          x = DIR < 0 ? zmul(x, w) : zmulc(x, w);
        }
        v[q][r] = x;
      }
      dft_small<R, DIR>(v[q], tw, L);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = threadIdx.x + q * nthr;
    if (j < nb) {
      const int k = j % Ns;
      const int o = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[o + r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// Lines l < nlines: element e of line l is at
//   (l / lpb) * bstride + (l % lpb) * lstride + e * estride.
// The first nin elements are read (the rest are zero), the first nout are
// written, times `scale`.  One workgroup per line.
struct LineArgs {
  const double2* in;
  double2* out;
  long nlines;
  long lpb, bstride, lstride, estride;
  int nin, nout;
  double scale;
};

template <int DIR>
__global__ void __launch_bounds__(1024) lines_kernel(LineArgs a, Plan P) {
  extern __shared__ double2 buf[];
  const long l = blockIdx.x;
  if (l >= a.nlines) return;
  const long base = (l / a.lpb) * a.bstride + (l % a.lpb) * a.lstride;
  const int L = P.L;
  const int nthr = blockDim.x;
  for (int e = threadIdx.x; e < L; e += nthr)
    buf[e] = e < a.nin ? a.in[base + e * a.estride] : make_double2(0.0, 0.0);
  __syncthreads();
  int Ns = 1;
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.radix[s];
    switch (R) {
      case 2: lds_stage<2, DIR>(buf, L, Ns, P.tw); break;
      case 3: lds_stage<3, DIR>(buf, L, Ns, P.tw); break;
      case 4: lds_stage<4, DIR>(buf, L, Ns, P.tw); break;
      case 5: lds_stage<5, DIR>(buf, L, Ns, P.tw); break;
      default: lds_stage<7, DIR>(buf, L, Ns, P.tw); break;
    }
    Ns *= R;
  }
  for (int e = threadIdx.x; e < a.nout; e += nthr) {
    const double2 x = buf[e];
    a.out[base + e * a.estride] = make_double2(x.x * a.scale, x.y * a.scale);
  }
}

// threads per line: every thread holds at most 4 butterflies of a stage
static int line_threads(int L) {
  int t = 64;
  while (t < 1024 && (L / 2 + t - 1) / t > 4) t *= 2;
  return t;
}

static void run_lines(const LineArgs& a, int L, int dir, hipStream_t s) {
  if (a.nlines <= 0) return;
  const Plan& P = length(L).plan();
  const size_t smem = sizeof(double2) * (size_t)L;
  const int nthr = line_threads(L);
  TC_REQUIRE(a.nlines < (1L << 31), "FP64 FFT: too many lines");
  auto launch = [&](auto k) {
    if (smem > 64 * 1024)
      TC_HIP(hipFuncSetAttribute((const void*)k,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)smem));
    hipLaunchKernelGGL(k, dim3((unsigned)a.nlines), dim3(nthr), smem, s, a,
                       P);
  };
  if (dir < 0)
    launch(lines_kernel<-1>);
  else
    launch(lines_kernel<1>);
  TC_HIP(hipGetLastError());
}

// in-place unnormalised 2D transform of nbatch [rows][cols] arrays
static void fft2(double2* d, int nbatch, int rows, int cols, int dir,
                 hipStream_t s) {
  LineArgs a{d, d, (long)nbatch * rows, (long)nbatch * rows, 0, cols, 1,
             cols, cols, 1.0};
  run_lines(a, cols, dir, s);
  LineArgs b{d, d, (long)nbatch * cols, cols, (long)rows * cols, 1, cols,
             rows, rows, 1.0};
  run_lines(b, rows, dir, s);
}

// ---------------------------------------------------------------------------
// small vector kernels
// ---------------------------------------------------------------------------

__global__ void zero_kernel(double2* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}
static void zero(double2* p, long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(zero_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n);
  TC_HIP(hipGetLastError());
}

// out = alpha a + beta b (b may be null: out = alpha a)
__global__ void axpby_kernel(const double* a, const double* b, double* out,
                             long n, double alpha, double beta) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = b ? alpha * a[i] + beta * b[i] : alpha * a[i];
}
static void axpby(const double* a, const double* b, double* out, long n,
                  double alpha, double beta, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, a, b,
                     out, n, alpha, beta);
  TC_HIP(hipGetLastError());
}

__global__ void fill_kernel(double* p, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

constexpr int kRedBlocks = 512;

__global__ void __launch_bounds__(256)
dot_kernel(const double* a, const double* b, long n, double* part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long)gridDim.x * 256)
    acc += a[i] * b[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sum of part[0..n) in a fixed order into part[n]
__global__ void sum_kernel(double* part, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += part[i];
    part[n] = t;
  }
}

static double reduce_host(double* part, int n, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, part, n);
  TC_HIP(hipGetLastError());
  double h = 0.0;
  TC_HIP(hipMemcpyAsync(&h, part + n, sizeof(double), hipMemcpyDeviceToHost,
                        s));
  TC_HIP(hipStreamSynchronize(s));
  return h;
}

double dot(const double* a, const double* b, long n, hipStream_t s) {
  DevBuf part(sizeof(double) * (kRedBlocks + 1));
  hipLaunchKernelGGL(dot_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, b, n,
                     part.as<double>());
  TC_HIP(hipGetLastError());
  return reduce_host(part.as<double>(), kRedBlocks, s);
}

__global__ void __launch_bounds__(256)
max_kernel(const double* a, long n, double* part) {
  __shared__ double red[256];
  double m = -1.0e300;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long)gridDim.x * 256)
    m = fmax(m, a[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

static void minmax(const double* a, long n, double* mx, double* mn,
                   hipStream_t s) {
  DevBuf part(sizeof(double) * 2 * kRedBlocks);
  DevBuf neg(sizeof(double) * n);
  hipLaunchKernelGGL(max_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, n,
                     part.as<double>());
  axpby(a, nullptr, neg.as<double>(), n, -1.0, 0.0, s);
  hipLaunchKernelGGL(max_kernel, dim3(kRedBlocks), dim3(256), 0, s,
                     neg.as<double>(), n, part.as<double>() + kRedBlocks);
  TC_HIP(hipGetLastError());
  std::vector<double> h(2 * kRedBlocks);
  TC_HIP(hipMemcpyAsync(h.data(), part.as<void>(),
                        sizeof(double) * 2 * kRedBlocks,
                        hipMemcpyDeviceToHost, s));
  TC_HIP(hipStreamSynchronize(s));
  double M = -1.0e300, m = -1.0e300;
  for (int i = 0; i < kRedBlocks; ++i) {
    M = std::max(M, h[i]);
    m = std::max(m, h[kRedBlocks + i]);
  }
  *mx = M;
  *mn = -m;
}

// ---------------------------------------------------------------------------
// NUFFT (exponential-of-semicircle kernel, upsampling 2)
// ---------------------------------------------------------------------------

constexpr int kW = 16;                  // kernel width (points)
constexpr double kBeta = 2.30 * kW;     // FINUFFT's beta for upsampling 2

__device__ __forceinline__ double es_phi(double z) {
  return fabs(z) <= 1.0 ? exp(kBeta * (sqrt(1.0 - z * z) - 1.0)) : 0.0;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
static void gauss_legendre(int n, std::vector<double>& x,
                           std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

Nufft::Nufft(const double* h_angles, int P, int N, int nmodes)
    : P_(P), N_(N), nm_(nmodes) {
  TC_REQUIRE(P >= 1 && N >= 2 && nmodes >= 1, "FP64 NUFFT: bad geometry");
  nf_ = smooth_at_least(std::max(2 * nmodes, 2 * kW));
  if (nf_ % 2) nf_ = smooth_at_least(nf_ + 1);
  TC_REQUIRE(nf_ <= kMaxFFT,
             "FP64 NUFFT: fine grid of " + std::to_string(nf_) +
                 " points exceeds " + std::to_string(kMaxFFT) +
                 " (widths up to 2047 after padding)");
  std::vector<double2> cs(P);
  for (int p = 0; p < P; ++p)
    cs[p] = make_double2(std::cos(h_angles[p]), std::sin(h_angles[p]));
  // p(k) = h / psi_hat(k), psi_hat(k) = (w h / 2) int phi(z) cos(k z w h/2)
  const double h = 2.0 * M_PI / nf_;
  std::vector<double> zq, wq;
  gauss_legendre(200, zq, wq);
  std::vector<double> corr(nmodes);
  for (int i = 0; i < nmodes; ++i) {
    const int k = i - nmodes / 2;
    double integ = 0.0;
    for (size_t q = 0; q < zq.size(); ++q) {
      const double z = zq[q];
      const double ph = std::exp(kBeta * (std::sqrt(1.0 - z * z) - 1.0));
      integ += wq[q] * ph * std::cos(k * z * kW * h / 2.0);
    }
    corr[i] = h / ((kW * h / 2.0) * integ) * ((k & 1) ? -1.0 : 1.0);
  }
  cs_.alloc(sizeof(double2) * P);
  corr_.alloc(sizeof(double) * nmodes);
  TC_HIP(hipMemcpy(cs_.as<void>(), cs.data(), sizeof(double2) * P,
                   hipMemcpyHostToDevice));
  TC_HIP(hipMemcpy(corr_.as<void>(), corr.data(), sizeof(double) * nmodes,
                   hipMemcpyHostToDevice));
}

struct GridArgs {
  const double2* cs;
  int P, N, nf, S;
};

// fine-grid coordinate of point (p, j) along x (c.x) or y (c.y)
__device__ __forceinline__ void point_uv(const GridArgs& g, int p, int j,
                                         double* u, double* v) {
  const double2 c = g.cs[p];
  const double rr = (double)(j - g.N / 2) * ((double)g.nf / (g.N - 1));
  *u = rr * c.x + 0.5 * g.nf;
  *v = rr * c.y + 0.5 * g.nf;
}

// type 1: every (slice, point) adds c_j phi phi to its w x w footprint
__global__ void spread_kernel(GridArgs g, const double2* __restrict__ c,
                              double2* fine) {
  const long J = (long)g.P * g.N;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J * g.S) return;
  const int s = (int)(i / J);
  const long jj = i - (long)s * J;
  const int p = (int)(jj / g.N), j = (int)(jj - (long)p * g.N);
  double u, v;
  point_uv(g, p, j, &u, &v);
  const double2 val = c[i];
  const int lx = (int)ceil(u - 0.5 * kW), ly = (int)ceil(v - 0.5 * kW);
  double wx[kW], wy[kW];
  for (int o = 0; o < kW; ++o) {
    wx[o] = es_phi((lx + o - u) * (2.0 / kW));
    wy[o] = es_phi((ly + o - v) * (2.0 / kW));
  }
  double* f = reinterpret_cast<double*>(fine + (long)s * g.nf * g.nf);
  for (int a = 0; a < kW; ++a) {
    int iy = (ly + a) % g.nf;
    if (iy < 0) iy += g.nf;
    const double wa = wy[a];
    for (int b = 0; b < kW; ++b) {
      int ix = (lx + b) % g.nf;
      if (ix < 0) ix += g.nf;
      const double wgt = wa * wx[b];
      const long o = 2 * ((long)iy * g.nf + ix);
      unsafeAtomicAdd(f + o, wgt * val.x);
      unsafeAtomicAdd(f + o + 1, wgt * val.y);
    }
  }
}

// type 2: every (slice, point) gathers its footprint
__global__ void interp_kernel(GridArgs g, const double2* __restrict__ fine,
                              double2* c) {
  const long J = (long)g.P * g.N;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J * g.S) return;
  const int s = (int)(i / J);
  const long jj = i - (long)s * J;
  const int p = (int)(jj / g.N), j = (int)(jj - (long)p * g.N);
  double u, v;
  point_uv(g, p, j, &u, &v);
  const int lx = (int)ceil(u - 0.5 * kW), ly = (int)ceil(v - 0.5 * kW);
  double wx[kW];
  for (int o = 0; o < kW; ++o) wx[o] = es_phi((lx + o - u) * (2.0 / kW));
  const double2* f = fine + (long)s * g.nf * g.nf;
  double2 acc = make_double2(0.0, 0.0);
  for (int a = 0; a < kW; ++a) {
    int iy = (ly + a) % g.nf;
    if (iy < 0) iy += g.nf;
    const double wa = es_phi((ly + a - v) * (2.0 / kW));
    double2 row = make_double2(0.0, 0.0);
    for (int b = 0; b < kW; ++b) {
      int ix = (lx + b) % g.nf;
      if (ix < 0) ix += g.nf;
      const double2 q = f[(long)iy * g.nf + ix];
      row.x += wx[b] * q.x;
      row.y += wx[b] * q.y;
    }
    acc.x += wa * row.x;
    acc.y += wa * row.y;
  }
  c[i] = acc;
}

// type 1 output: f[s][k2][k1] = corr(k1) corr(k2) fine[k2 mod nf][k1 mod nf]
__global__ void extract_kernel(const double2* __restrict__ fine,
                               const double* __restrict__ corr, int nm,
                               int nf, int S, double2* f, double* fr,
                               double rscale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nm * nm;
  if (i >= per * S) return;
  const int s = (int)(i / per);
  const int rem = (int)(i - (long)s * per);
  const int i2 = rem / nm, i1 = rem - i2 * nm;
  int m1 = (i1 - nm / 2) % nf, m2 = (i2 - nm / 2) % nf;
  if (m1 < 0) m1 += nf;
  if (m2 < 0) m2 += nf;
  const double2 q = fine[((long)s * nf + m2) * nf + m1];
  const double w = corr[i1] * corr[i2];
  if (fr)
    fr[i] = q.x * w * rscale;
  else
    f[i] = make_double2(q.x * w, q.y * w);
}

// type 2 input: fine[k2 mod nf][k1 mod nf] = corr corr f (fine zeroed)
__global__ void place_kernel(const double* __restrict__ f,
                             const double* __restrict__ corr, int nm, int nf,
                             int S, double2* fine) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nm * nm;
  if (i >= per * S) return;
  const int s = (int)(i / per);
  const int rem = (int)(i - (long)s * per);
  const int i2 = rem / nm, i1 = rem - i2 * nm;
  int m1 = (i1 - nm / 2) % nf, m2 = (i2 - nm / 2) % nf;
  if (m1 < 0) m1 += nf;
  if (m2 < 0) m2 += nf;
  fine[((long)s * nf + m2) * nf + m1] =
      make_double2(f[i] * corr[i1] * corr[i2], 0.0);
}

void Nufft::type1(const double2* c, double2* f, double* fr, double rscale,
                  int S, hipStream_t s) {
  const long fplane = (long)nf_ * nf_;
  fine_.reserve(sizeof(double2) * (size_t)S * fplane);
  double2* fine = fine_.as<double2>();
  zero(fine, S * fplane, s);
  GridArgs g{cs_.as<double2>(), P_, N_, nf_, S};
  const long npts = (long)S * P_ * N_;
  hipLaunchKernelGGL(spread_kernel, dim3(cdiv(npts, 256)), dim3(256), 0, s,
                     g, c, fine);
  TC_HIP(hipGetLastError());
  fft2(fine, S, nf_, nf_, +1, s);
  const long nout = (long)S * nm_ * nm_;
  hipLaunchKernelGGL(extract_kernel, dim3(cdiv(nout, 256)), dim3(256), 0, s,
                     fine, corr_.as<double>(), nm_, nf_, S, f, fr, rscale);
  TC_HIP(hipGetLastError());
}

void Nufft::type2_real(const double* f, double2* c, int S, hipStream_t s) {
  const long fplane = (long)nf_ * nf_;
  fine_.reserve(sizeof(double2) * (size_t)S * fplane);
  double2* fine = fine_.as<double2>();
  zero(fine, S * fplane, s);
  const long nin = (long)S * nm_ * nm_;
  hipLaunchKernelGGL(place_kernel, dim3(cdiv(nin, 256)), dim3(256), 0, s, f,
                     corr_.as<double>(), nm_, nf_, S, fine);
  TC_HIP(hipGetLastError());
  fft2(fine, S, nf_, nf_, -1, s);
  GridArgs g{cs_.as<double2>(), P_, N_, nf_, S};
  const long npts = (long)S * P_ * N_;
  hipLaunchKernelGGL(interp_kernel, dim3(cdiv(npts, 256)), dim3(256), 0, s, g,
                     fine, c);
  TC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// projectors: centred ray DFT fftshift(F(ifftshift v)) (src/gpu/fftshift.cu,
// src/fft.h) by direct summation: out[k] = sum_z v[z] w^(sign (z-c)(k-c)),
// c = N/2, w = exp(2 pi i / N)
// ---------------------------------------------------------------------------

template <bool IN_REAL, int SIGN, bool RAMP>
__global__ void ray_dft_kernel(const void* __restrict__ in, double2* out,
                               long nrays, int N,
                               const double2* __restrict__ tw) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrays * N) return;
  const long ray = i / N;
  const int k = (int)(i - ray * N);
  const int c = N / 2;
  const long kk = k - c;
  double2 acc = make_double2(0.0, 0.0);
  for (int z = 0; z < N; ++z) {
    long m = ((long)(z - c) * kk) % N;
    if (m < 0) m += N;
    const double2 w = tw[m];  // exp(-2 pi i m / N)
    double2 x;
    if constexpr (IN_REAL)
      x = make_double2(reinterpret_cast<const double*>(in)[ray * N + z], 0.0);
    else
      x = reinterpret_cast<const double2*>(in)[ray * N + z];
    const double2 t = SIGN < 0 ? zmul(x, w) : zmulc(x, w);
    acc.x += t.x;
    acc.y += t.y;
  }
  if (RAMP) {
    const double r = fabs((double)(k - c)) / N;  // |k - N/2| / N (< 1)
    acc.x *= r;
    acc.y *= r;
  }
  out[i] = acc;
}

__global__ void real_part_kernel(const double2* __restrict__ in, double* out,
                                 long n, double scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i].x * scale;
}

static std::shared_ptr<Nufft> nufft_plan(const double* h_angles, int P, int N,
                                         int nmodes) {
  // plans are cheap to build here (host quadrature, two small copies) and
  // own scratch, so each call builds its own
  ArenaPause no_arena;
  return std::make_shared<Nufft>(h_angles, P, N, nmodes);
}

void project(const double* img, double* sino, int S, int N,
             const double* h_angles, int P, hipStream_t s) {
  TC_REQUIRE(S >= 1 && N >= 2 && P >= 1, "FP64 project: bad shape");
  auto nu = nufft_plan(h_angles, P, N, N);
  const long ns = (long)S * P * N;
  DevBuf c(sizeof(double2) * ns), v(sizeof(double2) * ns);
  nu->type2_real(img, c.as<double2>(), S, s);
  const double2* tw = roots(N);
  hipLaunchKernelGGL((ray_dft_kernel<false, 1, false>), dim3(cdiv(ns, 256)),
                     dim3(256), 0, s, c.as<void>(), v.as<double2>(),
                     (long)S * P, N, tw);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(real_part_kernel, dim3(cdiv(ns, 256)), dim3(256), 0, s,
                     v.as<double2>(), sino, ns, 1.0 / ((double)N * N));
  TC_HIP(hipGetLastError());
}

void backproject(const double* sino, double* img, int S, int P, int N,
                 const double* h_angles, bool fbp, hipStream_t s) {
  TC_REQUIRE(S >= 1 && N >= 2 && P >= 1, "FP64 backproject: bad shape");
  auto nu = nufft_plan(h_angles, P, N, N);
  const long ns = (long)S * P * N;
  DevBuf v(sizeof(double2) * ns);
  const double2* tw = roots(N);
  if (fbp)
    hipLaunchKernelGGL((ray_dft_kernel<true, -1, true>), dim3(cdiv(ns, 256)),
                       dim3(256), 0, s, sino, v.as<double2>(), (long)S * P, N,
                       tw);
  else
    hipLaunchKernelGGL((ray_dft_kernel<true, -1, false>), dim3(cdiv(ns, 256)),
                       dim3(256), 0, s, sino, v.as<double2>(), (long)S * P, N,
                       tw);
  TC_HIP(hipGetLastError());
  nu->type1(v.as<double2>(), nullptr, img, 1.0 / ((double)N * N), S, s);
}

void gradient_nufft(const double* x, const double* b, double* g, int S, int N,
                    const double* h_angles, int P, hipStream_t s) {
  TC_REQUIRE(S >= 1 && N >= 2 && P >= 1, "FP64 gradient: bad shape");
  auto nu = nufft_plan(h_angles, P, N, N);
  const long ns = (long)S * P * N;
  DevBuf c(sizeof(double2) * ns);
  nu->type2_real(x, c.as<double2>(), S, s);
  nu->type1(c.as<double2>(), nullptr, g, 1.0 / ((double)N * N * N), S, s);
  if (b) axpby(g, b, g, (long)S * N * N, 1.0, -1.0, s);
}

// ---------------------------------------------------------------------------
// Toeplitz operator
// ---------------------------------------------------------------------------

__global__ void psf_place_kernel(const double* __restrict__ psf, int n, int M,
                                 double2* circ) {
  const int n1 = 2 * n - 1;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n1 * n1) return;
  const int a = (int)(i / n1), b = (int)(i - (long)a * n1);
  const int r = (a - (n - 1) + M) % M, c = (b - (n - 1) + M) % M;
  circ[(long)r * M + c] = make_double2(psf[i], 0.0);
}

Psf::Psf(const double* h_angles, int P, int n, hipStream_t s) : n_(n) {
  TC_REQUIRE(n >= 2 && P >= 1, "FP64 PSF: bad shape");
  M_ = smooth_at_least(2 * n - 1);
  TC_REQUIRE(M_ <= kMaxFFT, "FP64 PSF: width too large");
  const int n1 = 2 * n - 1;
  // PSF = Re(type 1 of ones) on (2n - 1)^2 modes (src/toeplitz.h:51-73)
  DevBuf ones(sizeof(double2) * (size_t)P * n), psf(sizeof(double) * n1 * n1);
  {
    std::vector<double2> h((size_t)P * n, make_double2(1.0, 0.0));
    TC_HIP(hipMemcpyAsync(ones.as<void>(), h.data(), sizeof(double2) * h.size(),
                          hipMemcpyHostToDevice, s));
    auto nu = nufft_plan(h_angles, P, n, n1);
    nu->type1(ones.as<double2>(), nullptr, psf.as<double>(), 1.0, 1, s);
    TC_HIP(hipStreamSynchronize(s));  // h leaves scope
  }
  spec_.alloc(sizeof(double2) * (size_t)M_ * M_);
  zero(spec_.as<double2>(), (long)M_ * M_, s);
  hipLaunchKernelGGL(psf_place_kernel, dim3(cdiv((long)n1 * n1, 256)),
                     dim3(256), 0, s, psf.as<double>(), n, M_,
                     spec_.as<double2>());
  TC_HIP(hipGetLastError());
  fft2(spec_.as<double2>(), 1, M_, M_, -1, s);
}

__global__ void pad_kernel(const double* __restrict__ x, int n, int M, int S,
                           double2* w) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)M * M;
  if (i >= per * S) return;
  const int s = (int)(i / per);
  const int rem = (int)(i - (long)s * per);
  const int r = rem / M, c = rem - r * M;
  w[i] = make_double2(r < n && c < n ? x[((long)s * n + r) * n + c] : 0.0,
                      0.0);
}

__global__ void spec_mul_kernel(double2* w, const double2* __restrict__ h,
                                long per, int S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * S) return;
  w[i] = zmul(w[i], h[i % per]);
}

// out = Re(w[0:n][0:n]) scale (- b)
__global__ void crop_kernel(const double2* __restrict__ w, int n, int M,
                            int S, double scale, const double* b,
                            double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)n * n;
  if (i >= per * S) return;
  const int s = (int)(i / per);
  const int rem = (int)(i - (long)s * per);
  const int r = rem / n, c = rem - r * n;
  const double v = w[((long)s * M + r) * M + c].x * scale;
  out[i] = b ? v - b[i] : v;
}

void Psf::apply(const double* x, const double* b, double* out, int S,
                hipStream_t s) {
  TC_REQUIRE(S >= 1, "FP64 Toeplitz: need at least one slice");
  const long per = (long)M_ * M_;
  // slices per group: about 1 GiB of complex workspace
  const int group = (int)std::max<long>(1, std::min<long>(S, (1L << 30) /
                                                                 (16 * per)));
  if ((size_t)group > work_slices_) {
    work_.alloc(sizeof(double2) * (size_t)group * per);
    work_slices_ = group;
  }
  double2* w = work_.as<double2>();
  const long img = (long)n_ * n_;
  const double scale = 1.0 / ((double)n_ * n_ * n_ * (double)M_ * M_);
  for (int s0 = 0; s0 < S; s0 += group) {
    const int ns = std::min(group, S - s0);
    hipLaunchKernelGGL(pad_kernel, dim3(cdiv(per * ns, 256)), dim3(256), 0, s,
                       x + s0 * img, n_, M_, ns, w);
    TC_HIP(hipGetLastError());
    fft2(w, ns, M_, M_, -1, s);
    hipLaunchKernelGGL(spec_mul_kernel, dim3(cdiv(per * ns, 256)), dim3(256),
                       0, s, w, spec_.as<double2>(), per, ns);
    TC_HIP(hipGetLastError());
    fft2(w, ns, M_, M_, +1, s);
    hipLaunchKernelGGL(crop_kernel, dim3(cdiv(img * ns, 256)), dim3(256), 0,
                       s, w, n_, M_, ns, scale, b ? b + s0 * img : nullptr,
                       out + s0 * img);
    TC_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// qGGMRF (weights and rho' of src/gpu/potential_function.cuh:38-108)
// ---------------------------------------------------------------------------

constexpr double kMrfQ = 2.0, kMrfC = 0.001;

__device__ __forceinline__ double weight26(int dz, int dy, int dx) {
  const int k = (dz != 0) + (dy != 0) + (dx != 0);
  return k == 1 ? 0.0523 : (k == 2 ? 0.0370 : (k == 3 ? 0.0302 : 0.0));
}

// rho'(d) = sign(d) |d|^(q-1) / (sigma^q (c + t)) (q - (q - p) t / (c + t)),
// t = |d|^(q-p) / sigma^(q-p); rho'(0) = 0
__device__ __forceinline__ double d_rho(double d, double sigma, double p) {
  if (d == 0.0) return 0.0;
  const double ad = fabs(d);
  const double t1 = pow(ad, kMrfQ - p) / pow(sigma, kMrfQ - p);
  const double t2 = pow(ad, kMrfQ - 1.0);
  const double t3 = kMrfC + t1;
  const double mag = (t2 / (t3 * pow(sigma, kMrfQ))) *
                     (kMrfQ - ((kMrfQ - p) * t1) / t3);
  return d > 0.0 ? mag : -mag;
}

// rho(d) = |d/s|^q / (c + |d/s|^(q-p))
__device__ __forceinline__ double rho(double d, double sigma, double p) {
  const double u = fabs(d) / sigma;
  return pow(u, kMrfQ) / (kMrfC + pow(u, kMrfQ - p));
}

// FREE = false: neighbours clamped into the volume (the reference stencil);
// true: pairs leaving the volume are dropped
template <bool FREE>
__global__ void qggmrf_grad_kernel(const double* __restrict__ x, double* g,
                                   int S, int R, int C, double sigma,
                                   double p) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)R * C;
  if (i >= per * S) return;
  const int s = (int)(i / per);
  const int rem = (int)(i - (long)s * per);
  const int r = rem / C, c = rem - r * C;
  const double v = x[i];
  double acc = 0.0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (!dz && !dy && !dx) continue;
        int z = s + dz, y = r + dy, xx = c + dx;
        if (FREE) {
          if (z < 0 || z >= S || y < 0 || y >= R || xx < 0 || xx >= C)
            continue;
        } else {
          z = min(max(z, 0), S - 1);
          y = min(max(y, 0), R - 1);
          xx = min(max(xx, 0), C - 1);
        }
        acc += weight26(dz, dy, dx) *
               d_rho(v - x[((long)z * R + y) * C + xx], sigma, p);
      }
  g[i] += acc;
}

// 1/2 sum over in-bounds ordered pairs of w rho (free boundary), per block
__global__ void __launch_bounds__(256)
qggmrf_value_kernel(const double* __restrict__ x, int S, int R, int C,
                    double sigma, double p, double* part) {
  __shared__ double red[256];
  const long per = (long)R * C;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < per * S;
       i += (long)gridDim.x * 256) {
    const int s = (int)(i / per);
    const int rem = (int)(i - (long)s * per);
    const int r = rem / C, c = rem - r * C;
    const double v = x[i];
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (!dz && !dy && !dx) continue;
          const int z = s + dz, y = r + dy, xx = c + dx;
          if (z < 0 || z >= S || y < 0 || y >= R || xx < 0 || xx >= C)
            continue;
          acc += weight26(dz, dy, dx) *
                 rho(v - x[((long)z * R + y) * C + xx], sigma, p);
        }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = 0.5 * red[0];
}

void add_total_var(const double* x, double* g, int S, int R, int C,
                   double sigma, double p, hipStream_t s) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1 && sigma > 0.0, "FP64 qGGMRF: bad "
                                                        "arguments");
  const long n = (long)S * R * C;
  hipLaunchKernelGGL(qggmrf_grad_kernel<false>, dim3(cdiv(n, 256)), dim3(256),
                     0, s, x, g, S, R, C, sigma, p);
  TC_HIP(hipGetLastError());
}

void add_qggmrf_grad_free(const double* x, double* g, int S, int R, int C,
                          double sigma, double p, hipStream_t s) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1 && sigma > 0.0, "FP64 qGGMRF: bad "
                                                        "arguments");
  const long n = (long)S * R * C;
  hipLaunchKernelGGL(qggmrf_grad_kernel<true>, dim3(cdiv(n, 256)), dim3(256),
                     0, s, x, g, S, R, C, sigma, p);
  TC_HIP(hipGetLastError());
}

double qggmrf_value_free(const double* x, int S, int R, int C, double sigma,
                         double p, hipStream_t s) {
  TC_REQUIRE(S >= 1 && R >= 1 && C >= 1 && sigma > 0.0, "FP64 qGGMRF: bad "
                                                        "arguments");
  DevBuf part(sizeof(double) * (kRedBlocks + 1));
  hipLaunchKernelGGL(qggmrf_value_kernel, dim3(kRedBlocks), dim3(256), 0, s,
                     x, S, R, C, sigma, p, part.as<double>());
  TC_HIP(hipGetLastError());
  return reduce_host(part.as<double>(), kRedBlocks, s);
}

// ---------------------------------------------------------------------------
// pre/post-processing (src/preproc.cpp:62-108, src/postproc.cpp:69-102,
// src/gpu/padding.cu:119-150)
// ---------------------------------------------------------------------------

__global__ void preproc_kernel(const double* __restrict__ in, double* out,
                               long rows, int N, int N2, int npad, int off,
                               double scale, double shift) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * N2) return;
  const long r = idx / N2;
  const int z = (int)(idx - r * N2);
  int zs = (z + off) % N2;
  if (zs < 0) zs += N2;
  const int zi = zs - npad;
  out[idx] = (zi >= 0 && zi < N) ? (in[r * N + zi] - shift) * scale : 0.0;
}

__global__ void pad2d_kernel(const double* __restrict__ in, double* out,
                             int S, int nin, int nout, int shift) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nout * nout;
  if (idx >= per * S) return;
  const long s = idx / per;
  const int rem = (int)(idx - s * per);
  const int r = rem / nout - shift, c = rem - (rem / nout) * nout - shift;
  out[idx] = (r >= 0 && r < nin && c >= 0 && c < nin)
                 ? in[(s * nin + r) * nin + c]
                 : 0.0;
}

__global__ void crop2d_kernel(const double* __restrict__ in, double* out,
                              int S, int nin, int nout, int shift) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nout * nout;
  if (idx >= per * S) return;
  const long s = idx / per;
  const int rem = (int)(idx - s * per);
  const int r = rem / nout, c = rem - r * nout;
  out[idx] = in[(s * nin + r + shift) * nin + c + shift];
}

static void pad_to(const double* in, double* out, int S, int nin, int nout,
                   hipStream_t s) {
  const long n = (long)S * nout * nout;
  hipLaunchKernelGGL(pad2d_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, in,
                     out, S, nin, nout, (nout - nin) / 2);
  TC_HIP(hipGetLastError());
}

static void crop2d(const double* in, double* out, int S, int nin, int nout,
                   hipStream_t s) {
  const long n = (long)S * nout * nout;
  hipLaunchKernelGGL(crop2d_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, in,
                     out, S, nin, nout, (nin - nout) / 2);
  TC_HIP(hipGetLastError());
}

// src/preproc.cpp:75-86 with the centre in FP64
static void pad_amount64(int nrays, double center, int* npad, int* offset) {
  const int cen = (int)std::round(center);  // halves away from zero
  const int off = (nrays + 1) / 2 - cen;
  int np = static_cast<int>(0.42 * nrays) / 2;
  if (std::abs(off) > np) np = std::abs(off);
  *npad = np;
  *offset = off;
}

int preproc(const double* in, double* out, int S, int P, int N, double center,
            hipStream_t s) {
  int npad = 0, off = 0;
  pad_amount64(N, center, &npad, &off);
  if (!out) return npad;
  TC_REQUIRE(in != nullptr, "null argument");
  const int n2 = N + 2 * npad;
  const long n = (long)S * P * n2;
  hipLaunchKernelGGL(preproc_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, in,
                     out, (long)S * P, N, n2, npad, off, 1.0, 0.0);
  TC_HIP(hipGetLastError());
  return npad;
}

void postproc(const double* in, double* out, int S, int nin, int nout,
              hipStream_t s) {
  TC_REQUIRE(nout >= 1 && nout <= nin, "postproc: bad shape");
  crop2d(in, out, S, nin, nout, s);
}

void pad2d(const double* in, double* out, int S, int nin, int padding,
           hipStream_t s) {
  TC_REQUIRE(S >= 1 && nin >= 1, "pad2d: bad shape");
  pad_to(in, out, S, nin, nin + std::abs(padding), s);
}

__global__ void __launch_bounds__(256)
diff2_kernel(const double* a, const double* b, long n, double* part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long)gridDim.x * 256) {
    const double d = a[i] - b[i];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

double diff_norm2(const double* a, const double* b, long n, hipStream_t s) {
  DevBuf part(sizeof(double) * (kRedBlocks + 1));
  hipLaunchKernelGGL(diff2_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, b, n,
                     part.as<double>());
  TC_HIP(hipGetLastError());
  return reduce_host(part.as<double>(), kRedBlocks, s);
}

// ---------------------------------------------------------------------------
// Optimizer::run2 (src/optimize.h:96-154) on device-resident FP64 state, with
// one operator application per attempt: gradient(y) = Ay - b + R'(y) for a
// linear A, and err(v) from a quadratic form (Toeplitz: ||y||^2 + v.Av -
// 2 v.b) or the projector (||Pv - y||)
// ---------------------------------------------------------------------------

struct Run2Ops {
  std::function<void(const double*, double*)> A;  // out = A v
  // err(v) given v and A v; for the projector form A v is unused
  std::function<double(const double*, const double*)> err;
};

static int run2(const Run2Ops& ops, const double* x0, const double* b,
                double* out, long N, int S, int R, int C, int num_iters,
                double step0, double sigma, double p, double* last_err,
                int* attempts, hipStream_t s) {
  DevBuf bs(sizeof(double) * N), bx(sizeof(double) * N),
      bTs(sizeof(double) * N), bTx(sizeof(double) * N),
      by(sizeof(double) * N), bTy(sizeof(double) * N), bg(sizeof(double) * N),
      bTg(sizeof(double) * N);
  double *sol = bs.as<double>(), *x = bx.as<double>(), *Ts = bTs.as<double>(),
         *Tx = bTx.as<double>(), *y = by.as<double>(), *Ty = bTy.as<double>(),
         *g = bg.as<double>(), *Tg = bTg.as<double>();
  TC_HIP(hipMemcpyAsync(sol, x0, sizeof(double) * N, hipMemcpyDeviceToDevice,
                        s));
  TC_HIP(hipMemcpyAsync(x, x0, sizeof(double) * N, hipMemcpyDeviceToDevice,
                        s));
  ops.A(sol, Ts);
  TC_HIP(hipMemcpyAsync(Tx, Ts, sizeof(double) * N, hipMemcpyDeviceToDevice,
                        s));
  double t = 1.0, tnew = 1.0, step = step0, fx = 0.0;
  int it = 0, att = 0;
  for (; it < num_iters; ++it) {
    for (;;) {
      const double beta = tnew * (1.0 / t - 1.0);
      tnew = 0.5 * (std::sqrt(t * t * t * t + 4.0 * t * t) - t * t);
      // y = sol + beta (sol - x), T y likewise (linearity)
      axpby(sol, x, y, N, 1.0 + beta, -beta, s);
      axpby(Ts, Tx, Ty, N, 1.0 + beta, -beta, s);
      // g = T y - b + R'(y)
      axpby(Ty, b, g, N, 1.0, -1.0, s);
      add_total_var(y, g, S, R, C, sigma, p, s);
      ops.A(g, Tg);
      // candidate and its image under A
      axpby(y, g, sol, N, 1.0, -step, s);
      axpby(Ty, Tg, Ts, N, 1.0, -step, s);
      ++att;
      fx = ops.err(sol, Ts);
      const double fy = ops.err(y, Ty);
      const double gy = 0.5 * step * dot(g, g, N, s);
      if (fx > fy + gy) {
        step *= 0.9;
        continue;
      }
      TC_HIP(hipMemcpyAsync(x, sol, sizeof(double) * N,
                            hipMemcpyDeviceToDevice, s));
      TC_HIP(hipMemcpyAsync(Tx, Ts, sizeof(double) * N,
                            hipMemcpyDeviceToDevice, s));
      step = step0;
      t = tnew;
      break;
    }
  }
  TC_HIP(hipMemcpyAsync(out, sol, sizeof(double) * N, hipMemcpyDeviceToDevice,
                        s));
  TC_HIP(hipStreamSynchronize(s));
  if (last_err) *last_err = fx;
  if (attempts) *attempts = att;
  return it;
}

Result mbir2(const double* sino, const double* x0, double* recon, int S,
             int P, int N, const double* h_angles, double center,
             int num_iters, double sigma, double p, hipStream_t s) {
  TC_REQUIRE(S >= 1 && P >= 1 && N >= 2, "FP64 mbir2: bad sinogram shape");
  TC_REQUIRE(h_angles != nullptr, "FP64 mbir2: null angles");
  TC_REQUIRE(num_iters >= 0, "FP64 mbir2: num_iters must be >= 0");
  ArenaScope arena;
  Result res;
  const long nsino = (long)S * P * N;
  const size_t img_b = sizeof(double) * (size_t)S * N * N;
  double mx = 0.0, mn = 0.0;
  minmax(sino, nsino, &mx, &mn, s);
  res.max_sino = mx;
  if (mx == 0.0) {
    if (!x0) throw Error(TC_ERR_VALUE, "mbir2: sinogram has ZERO max value");
    TC_HIP(hipMemcpyAsync(recon, x0, img_b, hipMemcpyDeviceToDevice, s));
    TC_HIP(hipStreamSynchronize(s));
    return res;
  }
  // normalise (src/mbir2.cpp:46-54) and the FBP guess (:56-60)
  DevBuf bs2(sizeof(double) * nsino);
  axpby(sino, nullptr, bs2.as<double>(), nsino, 1.0 / mx, 0.0, s);
  DevBuf bx0(img_b);
  if (x0)
    TC_HIP(hipMemcpyAsync(bx0.as<void>(), x0, img_b, hipMemcpyDeviceToDevice,
                          s));
  else
    backproject(bs2.as<double>(), bx0.as<double>(), S, P, N, h_angles, true,
                s);
  // pad and centre (:63-66)
  int npad = 0, off = 0;
  pad_amount64(N, center, &npad, &off);
  const int n2 = N + 2 * npad;
  res.padded = n2;
  const long ns2 = (long)S * P * n2;
  DevBuf bsp(sizeof(double) * ns2);
  hipLaunchKernelGGL(preproc_kernel, dim3(cdiv(ns2, 256)), dim3(256), 0, s,
                     bs2.as<double>(), bsp.as<double>(), (long)S * P, N, n2,
                     npad, off, 1.0, 0.0);
  TC_HIP(hipGetLastError());
  const long nimg = (long)S * n2 * n2;
  DevBuf bx0p(sizeof(double) * nimg), bb(sizeof(double) * nimg);
  pad_to(bx0.as<double>(), bx0p.as<double>(), S, N, n2, s);
  // b = P^T y, ||y||^2 (:74-77)
  backproject(bsp.as<double>(), bb.as<double>(), S, P, n2, h_angles, false,
              s);
  res.sino_norm = dot(bsp.as<double>(), bsp.as<double>(), ns2, s);
  // PSF (:86-97) and L = max(T 1) (:101-112)
  Psf psf(h_angles, P, n2, s);
  {
    const long nn = (long)n2 * n2;
    DevBuf ones(sizeof(double) * nn), t1(sizeof(double) * nn);
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, s,
                       ones.as<double>(), nn, 1.0);
    TC_HIP(hipGetLastError());
    psf.apply(ones.as<double>(), nullptr, t1.as<double>(), 1, s);
    double lmin = 0.0;
    minmax(t1.as<double>(), nn, &res.lipschitz, &lmin, s);
  }
  res.step = std::min(1.0, 1.0 / res.lipschitz);
  const double yd2 = res.sino_norm;
  const double* b = bb.as<double>();
  Run2Ops ops;
  ops.A = [&](const double* v, double* out) {
    psf.apply(v, nullptr, out, S, s);
  };
  // err(v) = sqrt(||y||^2 + v.Tv - 2 v.b) (src/error2.cpp:60-63,91)
  ops.err = [&](const double* v, const double* Tv) {
    return std::sqrt(yd2 + dot(v, Tv, nimg, s) - 2.0 * dot(v, b, nimg, s));
  };
  DevBuf bsol(sizeof(double) * nimg);
  res.iterations = run2(ops, bx0p.as<double>(), b, bsol.as<double>(), nimg, S,
                        n2, n2, num_iters, res.step, sigma, p,
                        &res.last_error, &res.attempts, s);
  crop2d(bsol.as<double>(), recon, S, n2, N, s);
  TC_HIP(hipStreamSynchronize(s));
  return res;
}

Result mbir_nufft(const double* sino, const double* x0, double* recon, int S,
                  int P, int N, const double* h_angles, double center,
                  int num_iters, double sigma, double p, hipStream_t s) {
  TC_REQUIRE(S >= 1 && P >= 1 && N >= 2, "FP64 mbir: bad sinogram shape");
  TC_REQUIRE(h_angles != nullptr && x0 != nullptr,
             "FP64 mbir: angles and an initial guess are required");
  TC_REQUIRE(num_iters >= 0, "FP64 mbir: num_iters must be >= 0");
  ArenaScope arena;
  Result res;
  const long nsino = (long)S * P * N;
  const size_t img_b = sizeof(double) * (size_t)S * N * N;
  double mx = 0.0, mn = 0.0;
  minmax(sino, nsino, &mx, &mn, s);
  res.max_sino = mx;
  if (mx == mn) {
    TC_HIP(hipMemcpyAsync(recon, x0, img_b, hipMemcpyDeviceToDevice, s));
    TC_HIP(hipStreamSynchronize(s));
    return res;
  }
  // normalise to [0, 1] and pad (src/mbir.cpp:43-60)
  int npad = 0, off = 0;
  pad_amount64(N, center, &npad, &off);
  const int n2 = N + 2 * npad;
  res.padded = n2;
  const long ns2 = (long)S * P * n2;
  DevBuf bsp(sizeof(double) * ns2);
  hipLaunchKernelGGL(preproc_kernel, dim3(cdiv(ns2, 256)), dim3(256), 0, s,
                     sino, bsp.as<double>(), (long)S * P, N, n2, npad, off,
                     1.0 / (mx - mn), mn);
  TC_HIP(hipGetLastError());
  const long nimg = (long)S * n2 * n2;
  DevBuf bx0p(sizeof(double) * nimg), bb(sizeof(double) * nimg);
  pad_to(x0, bx0p.as<double>(), S, N, n2, s);
  backproject(bsp.as<double>(), bb.as<double>(), S, P, n2, h_angles, false,
              s);
  res.sino_norm = dot(bsp.as<double>(), bsp.as<double>(), ns2, s);
  {
    const long nn = (long)n2 * n2;
    DevBuf ones(sizeof(double) * nn), t1(sizeof(double) * nn);
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, s,
                       ones.as<double>(), nn, 1.0);
    TC_HIP(hipGetLastError());
    gradient_nufft(ones.as<double>(), nullptr, t1.as<double>(), 1, n2,
                   h_angles, P, s);
    double lmin = 0.0;
    minmax(t1.as<double>(), nn, &res.lipschitz, &lmin, s);
  }
  res.step = std::min(1.0, 1.0 / res.lipschitz);
  DevBuf bproj(sizeof(double) * ns2);
  Run2Ops ops;
  ops.A = [&](const double* v, double* out) {
    gradient_nufft(v, nullptr, out, S, n2, h_angles, P, s);
  };
  // err(v) = ||P v - y|| (src/error.cpp:57-59)
  ops.err = [&](const double* v, const double*) {
    project(v, bproj.as<double>(), S, n2, h_angles, P, s);
    axpby(bproj.as<double>(), bsp.as<double>(), bproj.as<double>(), ns2, 1.0,
          -1.0, s);
    return std::sqrt(dot(bproj.as<double>(), bproj.as<double>(), ns2, s));
  };
  DevBuf bsol(sizeof(double) * nimg);
  res.iterations = run2(ops, bx0p.as<double>(), bb.as<double>(),
                        bsol.as<double>(), nimg, S, n2, n2, num_iters,
                        res.step, sigma, p, &res.last_error, &res.attempts, s);
  crop2d(bsol.as<double>(), recon, S, n2, N, s);
  TC_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace f64
}  // namespace tc
