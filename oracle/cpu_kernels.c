/*
 * CPU kernels of the FP32 CPU restatement of mbir2 (the benchmark's CPU
 * baseline).  Test and CPU-baseline infrastructure only: nothing in the
 * product links or loads this file.
 *
 * qggmrf_add_f32 restates cpuTotalVar of test/tvd.cpp:39-75 (weights,
 * d_potfun, replicate-at-edge clamp), parallelised over slices and rows with
 * OpenMP.  es_spread_f32 / es_interp_f32 are the spreading and interpolation
 * steps of an ES-kernel NUFFT on a centred fine grid (Barnett et al. 2019),
 * the same method as the GPU path.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const float W3[3][3][3] = {
    {{0.0302f, 0.037f, 0.0302f}, {0.037f, 0.0523f, 0.037f}, {0.0302f, 0.037f, 0.0302f}},
    {{0.037f, 0.0523f, 0.037f}, {0.0523f, 0.f, 0.0523f}, {0.037f, 0.0523f, 0.037f}},
    {{0.0302f, 0.037f, 0.0302f}, {0.037f, 0.0523f, 0.037f}, {0.0302f, 0.037f, 0.0302f}}};

/* test/tvd.cpp:24-36 */
static float d_potfun(float delta, float sigma, float p) {
  const float q = 2.f, c = 0.001f;
  float sigma_q = powf(sigma, q);
  float sigma_q_p = powf(sigma, q - p);
  float t1 = powf(fabsf(delta), q - p) / sigma_q_p;
  float t2 = powf(fabsf(delta), q - 1.f);
  float t3 = c + t1;
  if (delta > 0.f) return (t2 / (t3 * sigma_q)) * (q - ((q - p) * t1) / t3);
  if (delta < 0.f) return (-t2 / (t3 * sigma_q)) * (q - ((q - p) * t1) / t3);
  return 0.f;
}

/* g[v] += sum_d w_d rho'(x_v - x_clamp(v+d)) */
void qggmrf_add_f32(const float* x, float* g, int S, int R, int C, float sigma,
                    float p) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < S; ++i) {
    for (int j = 0; j < R; ++j) {
      for (int k = 0; k < C; ++k) {
        const float u = x[((size_t)i * R + j) * C + k];
        float v = 0.f;
        for (int z = 0; z < 3; ++z) {
          int i1 = i + z - 1;
          i1 = i1 < 0 ? 0 : (i1 >= S ? S - 1 : i1);
          for (int y = 0; y < 3; ++y) {
            int j1 = j + y - 1;
            j1 = j1 < 0 ? 0 : (j1 >= R ? R - 1 : j1);
            for (int xx = 0; xx < 3; ++xx) {
              int k1 = k + xx - 1;
              k1 = k1 < 0 ? 0 : (k1 >= C ? C - 1 : k1);
              float d = u - x[((size_t)i1 * R + j1) * C + k1];
              v += W3[z][y][xx] * d_potfun(d, sigma, p);
            }
          }
        }
        g[((size_t)i * R + j) * C + k] += v;
      }
    }
  }
}

static inline float es(float z, float beta) {
  float s = 1.f - z * z;
  return s >= 0.f ? expf(beta * (sqrtf(s) - 1.f)) : 0.f;
}

/* Type-1 spreading of c[nslices][J] (interleaved complex) at fine-grid
 * coordinates (u[J], v[J]) onto fine[nslices][nf][nf] (zeroed here).
 * Threads own disjoint bands of fine-grid rows, so no atomics are needed. */
void es_spread_f32(const float* c, const double* u, const double* v, long J,
                   int nslices, int nf, int w, float beta, float* fine) {
  memset(fine, 0, sizeof(float) * 2 * (size_t)nslices * nf * nf);
#pragma omp parallel
  {
#ifdef _OPENMP
    extern int omp_get_thread_num(void);
    extern int omp_get_num_threads(void);
    const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
#else
    const int tid = 0, nt = 1;
#endif
    const int r0 = (int)((long)nf * tid / nt), r1 = (int)((long)nf * (tid + 1) / nt);
    float wx[32], wy[32];
    for (long jj = 0; jj < J; ++jj) {
      const double hw = 0.5 * w;
      const int ly = (int)ceil(v[jj] - hw);
      /* quick reject: none of the w rows (mod nf) in [r0, r1) */
      int any = 0;
      for (int a = 0; a < w; ++a) {
        int r = ly + a;
        r += r < 0 ? nf : (r >= nf ? -nf : 0);
        if (r >= r0 && r < r1) { any = 1; break; }
      }
      if (!any) continue;
      const int lx = (int)ceil(u[jj] - hw);
      for (int a = 0; a < w; ++a) {
        wx[a] = es((float)((lx + a - u[jj]) * (2.0 / w)), beta);
        wy[a] = es((float)((ly + a - v[jj]) * (2.0 / w)), beta);
      }
      for (int a = 0; a < w; ++a) {
        int r = ly + a;
        r += r < 0 ? nf : (r >= nf ? -nf : 0);
        if (r < r0 || r >= r1) continue;
        for (int s = 0; s < nslices; ++s) {
          const float cr = c[2 * ((size_t)s * J + jj)];
          const float ci = c[2 * ((size_t)s * J + jj) + 1];
          float* row = fine + 2 * ((size_t)s * nf + r) * nf;
          for (int b = 0; b < w; ++b) {
            int q = lx + b;
            q += q < 0 ? nf : (q >= nf ? -nf : 0);
            const float wt = wy[a] * wx[b];
            row[2 * q] += cr * wt;
            row[2 * q + 1] += ci * wt;
          }
        }
      }
    }
  }
}

/* Type-2 interpolation: c[s][j] = sum fine[s][..] psi psi. */
void es_interp_f32(const float* fine, const double* u, const double* v, long J,
                   int nslices, int nf, int w, float beta, float* c) {
#pragma omp parallel for schedule(static)
  for (long jj = 0; jj < J; ++jj) {
    float wx[32], wy[32];
    int ix[32], iy[32];
    const double hw = 0.5 * w;
    const int lx = (int)ceil(u[jj] - hw), ly = (int)ceil(v[jj] - hw);
    for (int a = 0; a < w; ++a) {
      wx[a] = es((float)((lx + a - u[jj]) * (2.0 / w)), beta);
      wy[a] = es((float)((ly + a - v[jj]) * (2.0 / w)), beta);
      int qx = lx + a, qy = ly + a;
      ix[a] = qx + (qx < 0 ? nf : (qx >= nf ? -nf : 0));
      iy[a] = qy + (qy < 0 ? nf : (qy >= nf ? -nf : 0));
    }
    for (int s = 0; s < nslices; ++s) {
      float ar = 0.f, ai = 0.f;
      for (int a = 0; a < w; ++a) {
        const float* row = fine + 2 * ((size_t)s * nf + iy[a]) * nf;
        float rr = 0.f, ri = 0.f;
        for (int b = 0; b < w; ++b) {
          rr += row[2 * ix[b]] * wx[b];
          ri += row[2 * ix[b] + 1] * wx[b];
        }
        ar += rr * wy[a];
        ai += ri * wy[a];
      }
      c[2 * ((size_t)s * J + jj)] = ar;
      c[2 * ((size_t)s * J + jj) + 1] = ai;
    }
  }
}
