// extern "C" entry points (declared in include/tomocam_hip.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "driver.h"
#include "mbir.h"
#include "nufft.h"
#include "toeplitz.h"
#include "vecops.h"

namespace tc {
void ktimer_enable(int kc, bool on);
void ktimer_set_stride(int stride);
void ktimer_collect(int kc, double* total_ms, long* count, double* bytes);
}  // namespace tc

struct tc_psf_s {
  std::unique_ptr<tc::Psf> psf;
};

using tc::guarded;

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static tc::Psf& get_psf(tc_psf p) {
  TC_REQUIRE(p != nullptr && p->psf, "null PSF handle");
  return *p->psf;
}

extern "C" {

const char* tc_last_error(void) { return tc::last_error(); }

int tc_version(int* major, int* minor, int* patch) {
  return guarded([&] {
    if (major) *major = 0;
    if (minor) *minor = 1;
    if (patch) *patch = 0;
  });
}

int tc_set_num_of_gpus(int n) {
  return guarded([&] {
    TC_REQUIRE(n >= 1, "set_num_of_gpus: n must be >= 1");
  });
}

int tc_project(const float* d_img, float* d_sino, int nslices, int npix,
               const float* h_angles, int nproj, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_img && d_sino && h_angles, "null argument");
    // temporaries from the thread's arena, reused in stream order: no
    // hipMalloc / hipFree and no synchronisation per call
    tc::ArenaScope arena(as_stream(stream));
    tc::project(d_img, d_sino, nslices, npix, h_angles, nproj,
                as_stream(stream));
  });
}

int tc_backproject(const float* d_sino, float* d_img, int nslices, int nproj,
                   int npix, const float* h_angles, int fbp, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_img && d_sino && h_angles, "null argument");
    tc::ArenaScope arena(as_stream(stream));
    tc::backproject(d_sino, d_img, nslices, nproj, npix, h_angles, fbp != 0,
                    as_stream(stream));
  });
}

int tc_gradient(const float* d_x, const float* d_sinoT, float* d_g,
                int nslices, int npix, const float* h_angles, int nproj,
                void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && d_sinoT && d_g && h_angles, "null argument");
    TC_REQUIRE(nslices >= 1 && npix >= 2 && nproj >= 1, "gradient: bad shape");
    hipStream_t s = as_stream(stream);
    tc::ArenaScope arena(s);
    tc::PolarNufft& nu = tc::nufft_plan(h_angles, nproj, npix, npix);
    tc::DevBuf c(sizeof(float2) * (size_t)nslices * nproj * npix);
    nu.type2_real(d_x, c.as<float2>(), nslices, s);
    const double n3 = (double)npix * npix * npix;
    nu.type1_real(c.as<float2>(), d_g, nslices, (float)(1.0 / n3), s);
    const long N = (long)nslices * npix * npix;
    tc::affine_sub(d_g, d_sinoT, d_g, N, s);
  });
}

int tc_function_value(const float* d_x, const float* d_sino, int nslices,
                      int npix, const float* h_angles, int nproj,
                      double* h_out, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && d_sino && h_angles && h_out, "null argument");
    tc::ArenaScope arena;
    hipStream_t s = as_stream(stream);
    const long ns = (long)nslices * nproj * npix;
    tc::DevBuf px(sizeof(float) * ns);
    tc::project(d_x, px.as<float>(), nslices, npix, h_angles, nproj, s);
    *h_out = tc::diff_norm2(px.as<float>(), d_sino, ns, s);
  });
}

int tc_xerror(const float* d_a, const float* d_b, long n, double* h_out,
              void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_a && d_b && h_out, "null argument");
    *h_out = tc::diff_norm2(d_a, d_b, n, as_stream(stream));
  });
}

int tc_preproc(const float* d_sino, float* d_out, int nslices, int nproj,
               int nrays, float center, int* h_npad, void* stream) {
  return guarded([&] {
    TC_REQUIRE(nslices >= 1 && nproj >= 1 && nrays >= 1, "preproc: bad shape");
    int npad = 0, off = 0;
    tc::pad_amount(nrays, center, &npad, &off);
    if (h_npad) *h_npad = npad;
    if (d_out) {
      TC_REQUIRE(d_sino != nullptr, "null argument");
      tc::preproc(d_sino, d_out, nslices, nproj, nrays, npad, off,
                  as_stream(stream));
    }
  });
}

int tc_postproc(const float* d_img, float* d_out, int nslices, int n,
                int npixels, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_img && d_out, "null argument");
    TC_REQUIRE(nslices >= 1 && npixels >= 1 && npixels <= n, "postproc: bad shape");
    tc::postproc(d_img, d_out, nslices, n, npixels, as_stream(stream));
  });
}

int tc_pad2d(const float* d_img, float* d_out, int nslices, int n,
             int padding, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_img && d_out, "null argument");
    TC_REQUIRE(nslices >= 1 && n >= 1, "pad2d: bad shape");
    tc::pad2d(d_img, d_out, nslices, n, padding, as_stream(stream));
  });
}

int tc_psf_create(tc_psf* out, const float* h_angles, int nproj, int npix,
                  void* stream) {
  return guarded([&] {
    TC_REQUIRE(out != nullptr && h_angles != nullptr, "null argument");
    TC_REQUIRE(nproj >= 1 && npix >= 2, "psf: bad geometry");
    hipStream_t s = as_stream(stream);
    const size_t n1 = 2 * (size_t)npix - 1;
    tc::DevBuf arr(sizeof(float) * n1 * n1);
    tc::polar_psf(h_angles, nproj, npix, arr.as<float>(), s);
    auto h = std::make_unique<tc_psf_s>();
    h->psf = std::make_unique<tc::Psf>(arr.as<float>(), npix, s);
    *out = h.release();
  });
}

int tc_psf_create_from_array(tc_psf* out, const float* d_psf, int n,
                             void* stream) {
  return guarded([&] {
    TC_REQUIRE(out != nullptr && d_psf != nullptr, "null argument");
    auto h = std::make_unique<tc_psf_s>();
    h->psf = std::make_unique<tc::Psf>(d_psf, n, as_stream(stream));
    *out = h.release();
  });
}

int tc_psf_get_array(tc_psf p, float* d_psf, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    const size_t n1 = 2 * (size_t)P.n() - 1;
    TC_HIP(hipMemcpyAsync(d_psf, P.array(), sizeof(float) * n1 * n1,
                          hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}

int tc_psf_destroy(tc_psf p) {
  return guarded([&] { delete p; });
}

int tc_psf_info(tc_psf p, int* n, int* fft_size) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    if (n) *n = P.n();
    if (fft_size) *fft_size = P.M();
  });
}

int tc_toeplitz_apply(tc_psf p, const float* d_x, float* d_out, int nslices,
                      double* h_xTx, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x && d_out, "null argument");
    hipStream_t s = as_stream(stream);
    if (!h_xTx) {
      P.apply(d_x, nullptr, d_out, nslices, 0, nullptr, s);
      return;
    }
    const int np = tc::toeplitz_num_partials(P, nslices);
    tc::DevBuf parts(sizeof(double) * (np + 1));
    P.apply(d_x, nullptr, d_out, nslices, 0, parts.as<double>(), s);
    tc::reduce_partials(parts.as<double>(), np, parts.as<double>() + np, s);
    TC_HIP(hipMemcpyAsync(h_xTx, parts.as<double>() + np, sizeof(double),
                          hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
  });
}

// ---------------------------------------------------------------------------
// tc_toeplitz_apply_noalloc: tc_toeplitz_apply with its workspace (the
// row-transformed intermediate Wt of up to kScratchBytes / (8 H rp) slices
// at a time, the dot partials and their sum) in a statically allocated
// device array instead of library allocations.  The array exists once per
// GPU for the life of the process; host code reaches it only through the
// device address hipGetSymbolAddress returns on that GPU.
// ---------------------------------------------------------------------------

constexpr size_t kScratchBytes = size_t(72) << 20;  // 72 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device, looked up on the first
// call there.
void* scratch_on_current_device() {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  TC_REQUIRE(dev >= 0 && dev < kMaxDevices, "no-alloc: device index " +
                                                std::to_string(dev) +
                                                " out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TC_REQUIRE(p != nullptr, "no-alloc: scratch has no device address");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

int tc_toeplitz_apply_noalloc(tc_psf p, const float* d_x, float* d_out,
                              int nslices, double* h_xTx, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x && d_out, "null argument");
    TC_REQUIRE(nslices >= 1, "Toeplitz: need at least one slice");
    hipStream_t s = as_stream(stream);
    // layout: [dot partials (np) | their sum (1) | pad to 256 B | Wt]
    const size_t np = h_xTx ? (size_t)tc::toeplitz_num_partials(P, nslices)
                            : 0;
    const size_t head = (sizeof(double) * (np + 1) + 255) / 256 * 256;
    const size_t per = P.workspace_bytes_per_slice();
    if (head + per > kScratchBytes)
      throw tc::Error(TC_ERR_INVALID,
                      "no-alloc Toeplitz apply: needs " +
                          std::to_string(head + per) +
                          " bytes of scratch (one slice of width " +
                          std::to_string(P.n()) + " plus " +
                          std::to_string(np) + " partial sums), has " +
                          std::to_string(kScratchBytes));
    unsigned char* base = static_cast<unsigned char*>(
        scratch_on_current_device());
    double* parts = reinterpret_cast<double*>(base);
    P.apply_ws(d_x, d_out, nslices, h_xTx ? parts : nullptr, base + head,
               kScratchBytes - head, s);
    if (h_xTx) {
      // every partial is written by the column pass: nothing to clear
      tc::reduce_partials(parts, (long)np, parts + np, s);
      TC_HIP(hipMemcpyAsync(h_xTx, parts + np, sizeof(double),
                            hipMemcpyDeviceToHost, s));
      TC_HIP(hipStreamSynchronize(s));
    }
  });
}

int tc_toeplitz_fft_length(int n, int* fft_size) {
  return guarded([&] {
    TC_REQUIRE(n >= 2 && fft_size, "Toeplitz: bad width");
    *fft_size = tc::toeplitz_fft_length(n);
  });
}

int tc_psf_tune(tc_psf p, int nslices, int mixed, int* group, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(nslices >= 1, "Toeplitz: need at least one slice");
    const int g = P.tune_groups(nslices, mixed != 0, as_stream(stream));
    if (group) *group = g;
  });
}

int tc_psf_tune_length(tc_psf p, int nslices, int* fft_size, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(nslices >= 1, "Toeplitz: need at least one slice");
    P.tune_length(nslices, as_stream(stream));
    if (fft_size) *fft_size = P.M();
  });
}

int tc_toeplitz_apply_mixed(tc_psf p, const float* d_x, float* d_out,
                            int nslices, double* h_xTx, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x && d_out, "null argument");
    TC_REQUIRE(nslices >= 1, "Toeplitz: need at least one slice");
    TC_REQUIRE(P.has_fused(), "mixed precision needs a compile-time FFT "
                              "plan; FFT length " + std::to_string(P.M()) +
                              " has none");
    hipStream_t s = as_stream(stream);
    tc::DevBuf x16(tc::half_pair_bytes(nslices, P.n(), P.n()));
    tc::DevBuf gsc(sizeof(float) * (size_t)nslices);
    tc::to_half_slices(d_x, x16.as<void>(), gsc.as<float>(), nslices, P.n(),
                       P.n(), s);
    const int np = tc::toeplitz_num_partials(P, nslices);
    tc::DevBuf parts(sizeof(double) * (np + 1));
    P.apply_half(x16.as<void>(), gsc.as<float>(), d_out, nslices,
                 h_xTx ? parts.as<double>() : nullptr, s);
    if (h_xTx) {
      tc::reduce_partials(parts.as<double>(), np, parts.as<double>() + np, s);
      TC_HIP(hipMemcpyAsync(h_xTx, parts.as<double>() + np, sizeof(double),
                            hipMemcpyDeviceToHost, s));
    }
    TC_HIP(hipStreamSynchronize(s));
  });
}

int tc_gradient2(tc_psf p, const float* d_x, const float* d_sinoT, float* d_g,
                 int nslices, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x && d_sinoT && d_g, "null argument");
    P.apply(d_x, d_sinoT, d_g, nslices, 1, nullptr, as_stream(stream));
  });
}

int tc_function_value2(tc_psf p, const float* d_x, const float* d_sinoT,
                       int nslices, double sino_sq, double* h_out,
                       void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x && d_sinoT && h_out, "null argument");
    hipStream_t s = as_stream(stream);
    const long N = (long)nslices * P.n() * P.n();
    tc::DevBuf tx(sizeof(float) * N);
    const int np = tc::toeplitz_num_partials(P, nslices);
    tc::DevBuf parts(sizeof(double) * (np + 1));
    P.apply(d_x, nullptr, tx.as<float>(), nslices, 0, parts.as<double>(), s);
    tc::reduce_partials(parts.as<double>(), np, parts.as<double>() + np, s);
    double xtx = 0.0;
    TC_HIP(hipMemcpyAsync(&xtx, parts.as<double>() + np, sizeof(double),
                          hipMemcpyDeviceToHost, s));
    const double xb = tc::dot(d_x, d_sinoT, N, s);  // synchronises
    *h_out = sino_sq + xtx - 2.0 * xb;
  });
}

int tc_add_total_var2(const float* d_x, float* d_g, int nslices, int nrows,
                      int ncols, float sigma, float p, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && d_g, "null argument");
    tc::add_total_var(d_x, d_g, nslices, nrows, ncols, sigma, p,
                      as_stream(stream));
  });
}

int tc_add_total_var(const float* d_x, float* d_g, int nslices, int nrows,
                     int ncols, float sigma, float p, void* stream) {
  return tc_add_total_var2(d_x, d_g, nslices, nrows, ncols, sigma, p, stream);
}

int tc_add_tv_hessian(float* h_g, long n, float sigma, float* h_value) {
  return guarded([&] {
    (void)h_g;
    (void)n;
    TC_REQUIRE(sigma > 0.f, "add_tv_hessian: sigma must be > 0");
    if (h_value) *h_value = tc::tv_hessian_zero(sigma);
  });
}

int tc_qggmrf_grad_free(const float* d_x, float* d_g, int nslices, int nrows,
                        int ncols, float sigma, float p, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && d_g, "null argument");
    tc::add_qggmrf_grad_free(d_x, d_g, nslices, nrows, ncols, sigma, p,
                             as_stream(stream));
  });
}

int tc_qggmrf_value_free(const float* d_x, int nslices, int nrows, int ncols,
                         float sigma, float p, double* h_value, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && h_value, "null argument");
    tc::ArenaScope arena;  // synchronises before its temporaries return
    *h_value = tc::qggmrf_value_free(d_x, nslices, nrows, ncols, sigma, p,
                                     as_stream(stream));
  });
}

int tc_qggmrf_value_free_dev(const float* d_x, int nslices, int nrows,
                             int ncols, float sigma, float p, double* d_value,
                             void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_x && d_value, "null argument");
    TC_REQUIRE(sigma > 0.f, "qGGMRF: sigma must be > 0");
    tc::ArenaScope arena(as_stream(stream));
    tc::qggmrf_value_free_dev(d_x, nslices, nrows, ncols, sigma, p, d_value,
                              as_stream(stream));
  });
}

int tc_run2(tc_psf p, const float* d_x0, const float* d_sinoT, float* d_out,
            int nslices, int num_iters, double step, float sigma, float p_exp,
            double sino_sq, int refresh_every, int verbose, double* h_err,
            double* h_xerr, int* h_stats, void* stream) {
  return guarded([&] {
    tc::Psf& P = get_psf(p);
    TC_REQUIRE(d_x0 && d_sinoT && d_out, "null argument");
    tc::Run2Options o;
    o.num_iters = num_iters;
    o.step = step;
    o.sigma = sigma;
    o.p = p_exp;
    o.sino_norm = sino_sq;
    o.refresh_every = refresh_every;
    o.verbose = verbose;
    if (h_err || h_xerr) {
      o.on_iter = [&](int it, double e, double xe, double, int) {
        if (h_err) h_err[it] = e;
        if (h_xerr) h_xerr[it] = xe;
      };
    }
    tc::Run2Stats st =
        tc::run2_device(P, d_x0, d_sinoT, d_out, nslices, o, as_stream(stream));
    if (h_stats) {
      h_stats[0] = st.iterations;
      h_stats[1] = st.attempts;
      h_stats[2] = st.backtracks;
      h_stats[3] = st.toeplitz_applies;
    }
  });
}

static void fill_info(const tc::Mbir2Result& r, double* h_info) {
  if (!h_info) return;
  h_info[0] = r.lipschitz;
  h_info[1] = r.step;
  h_info[2] = r.sino_norm;
  h_info[3] = r.max_sino;
  h_info[4] = r.padded;
  h_info[5] = r.stats.iterations;
}

int tc_mbir2(const float* d_sino, const float* d_x0, float* d_recon,
             int nslices, int nproj, int nrays, const float* h_angles,
             float center, int num_iters, float sigma, float tol, float xtol,
             int refresh_every, int verbose, double* h_info, void* stream) {
  return guarded([&] {
    (void)tol;
    (void)xtol;
    TC_REQUIRE(d_sino && d_recon && h_angles, "null argument");
    TC_REQUIRE(num_iters >= 0, "mbir: num_iters must be >= 0");
    tc::Mbir2Options o;
    o.num_iters = num_iters;
    o.sigma = sigma;
    o.refresh_every = refresh_every;
    o.verbose = verbose;
    tc::Mbir2Result r = tc::mbir2(d_sino, d_x0, d_recon, nslices, nproj,
                                  nrays, h_angles, center, o,
                                  as_stream(stream));
    fill_info(r, h_info);
  });
}

int tc_mbir2_mixed(const float* d_sino, const float* d_x0, float* d_recon,
                   int nslices, int nproj, int nrays, const float* h_angles,
                   float center, int num_iters, float sigma,
                   int refresh_every, int lipschitz_iters, int verbose,
                   double* h_info, void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_sino && d_recon && h_angles, "null argument");
    TC_REQUIRE(num_iters >= 0, "mbir: num_iters must be >= 0");
    TC_REQUIRE(refresh_every >= 0, "mbir: refresh_every must be >= 0");
    TC_REQUIRE(lipschitz_iters >= 0, "mbir: lipschitz_iters must be >= 0");
    tc::Mbir2Options o;
    o.num_iters = num_iters;
    o.sigma = sigma;
    o.refresh_every = refresh_every;
    o.verbose = verbose;
    o.mixed = true;
    o.lipschitz_iters = lipschitz_iters;
    tc::Mbir2Result r = tc::mbir2(d_sino, d_x0, d_recon, nslices, nproj,
                                  nrays, h_angles, center, o,
                                  as_stream(stream));
    fill_info(r, h_info);
    if (h_info) {
      h_info[6] = r.stats.saturated;
      h_info[7] = r.stats.backtracks;
    }
  });
}

int tc_mbir(const float* d_sino, const float* d_x0, float* d_recon,
            int nslices, int nproj, int nrays, const float* h_angles,
            float center, int num_iters, float sigma, float tol, float xtol,
            int verbose, double* h_info, void* stream) {
  return guarded([&] {
    (void)tol;
    (void)xtol;
    TC_REQUIRE(d_sino && d_x0 && d_recon && h_angles, "null argument");
    TC_REQUIRE(num_iters >= 0, "mbir: num_iters must be >= 0");
    tc::Mbir2Options o;
    o.num_iters = num_iters;
    o.sigma = sigma;
    o.verbose = verbose;
    tc::Mbir2Result r = tc::mbir_nufft(d_sino, d_x0, d_recon, nslices, nproj,
                                       nrays, h_angles, center, o,
                                       as_stream(stream));
    fill_info(r, h_info);
  });
}

int tc_mbir2_batch(int nstacks, const float* const* d_sinos,
                   float* const* d_recons, const int* nslices,
                   const int* nproj, const int* nrays,
                   const float* const* h_angles, const float* centers,
                   int num_iters, float sigma, int refresh_every,
                   double* h_info, void* stream) {
  return guarded([&] {
    TC_REQUIRE(nstacks >= 0, "mbir batch: nstacks must be >= 0");
    TC_REQUIRE(nstacks == 0 || (d_sinos && d_recons && nslices && nproj &&
                                nrays && h_angles && centers),
               "null argument");
    TC_REQUIRE(num_iters >= 0, "mbir: num_iters must be >= 0");
    tc::Mbir2Options o;
    o.num_iters = num_iters;
    o.sigma = sigma;
    o.refresh_every = refresh_every;
    auto run_one = [&](int k, hipStream_t s) {
      TC_REQUIRE(d_sinos[k] && d_recons[k] && h_angles[k], "null argument");
      tc::Mbir2Result r = tc::mbir2(d_sinos[k], nullptr, d_recons[k],
                                    nslices[k], nproj[k], nrays[k],
                                    h_angles[k], centers[k], o, s);
      fill_info(r, h_info ? h_info + 6 * (long)k : nullptr);
    };
    // Stacks run on `workers` streams at once (TOMOCAM_HIP_BATCH_STREAMS,
    // default 4), one host thread each, largest stacks first: a small stack
    // alone leaves most of the GPU idle between its line-search
    // synchronisations.
    const char* ev = std::getenv("TOMOCAM_HIP_BATCH_STREAMS");
    int workers = ev ? std::atoi(ev) : 4;
    workers = std::max(1, std::min(workers, nstacks));
    if (workers == 1) {
      for (int k = 0; k < nstacks; ++k) {
        try {
          run_one(k, as_stream(stream));
        } catch (const tc::Error& e) {
          throw tc::Error(e.code, "stack " + std::to_string(k) + ": " +
                                        e.what());
        }
      }
      return;
    }
    TC_HIP(hipStreamSynchronize(as_stream(stream)));  // inputs are ready
    // timings taken while other stacks run would be noise: concurrent
    // stacks use tuned sizes from earlier calls, or the default rule
    o.autotune = false;
    // a worker capturing a graph would make the others' synchronous copies
    // (plan set-up) fail
    o.allow_graphs = false;
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    std::vector<int> order(nstacks);
    for (int k = 0; k < nstacks; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (double)nslices[a] * nrays[a] * nrays[a] >
             (double)nslices[b] * nrays[b] * nrays[b];
    });
    std::atomic<int> next{0};
    std::atomic<bool> stop{false};
    std::mutex emu;
    int err_stack = -1, err_code = TC_OK;
    std::string err_msg;
    auto fail = [&](int k, int code, const std::string& msg) {
      std::lock_guard<std::mutex> lk(emu);
      if (err_code == TC_OK || (k >= 0 && (err_stack < 0 || k < err_stack))) {
        err_stack = k;
        err_code = code;
        err_msg = msg;
      }
      stop = true;
    };
    auto work = [&](int /*worker*/) {
      int k = -1;
      try {
        TC_HIP(hipSetDevice(dev));
        hipStream_t s = nullptr;
        TC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamFree {
          hipStream_t s;
          ~StreamFree() { (void)hipStreamDestroy(s); }
        } sf{s};
        while (!stop) {
          const int i = next.fetch_add(1);
          if (i >= nstacks) break;
          k = order[i];
          run_one(k, s);
        }
      } catch (const tc::Error& e) {
        fail(k, e.code, e.what());
      } catch (const std::bad_alloc&) {
        fail(k, TC_ERR_NOMEM, "host allocation failed");
      } catch (const std::exception& e) {
        fail(k, TC_ERR_INTERNAL, e.what());
      }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(work, w);
    for (auto& t : pool) t.join();
    if (err_code != TC_OK)
      throw tc::Error(err_code, (err_stack >= 0 ? "stack " +
                                                       std::to_string(err_stack) +
                                                       ": "
                                                 : std::string()) +
                                    err_msg);
  });
}

int tc_projections_to_sinogram(const float* d_proj, float* d_sino, int nproj,
                               int nslices, int ncols, int drop_last_col,
                               void* stream) {
  return guarded([&] {
    TC_REQUIRE(d_proj && d_sino, "null argument");
    TC_REQUIRE(d_proj != d_sino, "projections: in and out must differ");
    tc::projections_to_sinogram(d_proj, d_sino, nproj, nslices, ncols,
                                drop_last_col != 0, as_stream(stream));
  });
}

int tc_mbir2_host(const float* h_sino, float* h_recon, int nslices, int nproj,
                  int nrays, const float* h_angles, float center,
                  int num_iters, float sigma, float tol, float xtol,
                  int verbose) {
  return guarded([&] {
    (void)tol;
    (void)xtol;
    TC_REQUIRE(h_sino && h_recon && h_angles, "null argument");
    TC_REQUIRE(nslices >= 1 && nproj >= 1 && nrays >= 2, "mbir: bad shape");
    const size_t ns = (size_t)nslices * nproj * nrays;
    const size_t nr = (size_t)nslices * nrays * nrays;
    tc::DevBuf ds(sizeof(float) * ns), dr(sizeof(float) * nr);
    TC_HIP(hipMemcpy(ds.as<void>(), h_sino, sizeof(float) * ns,
                     hipMemcpyHostToDevice));
    tc::Mbir2Options o;
    o.num_iters = num_iters;
    o.sigma = sigma;
    o.verbose = verbose;
    tc::mbir2(ds.as<float>(), nullptr, dr.as<float>(), nslices, nproj, nrays,
              h_angles, center, o, nullptr);
    TC_HIP(hipMemcpy(h_recon, dr.as<void>(), sizeof(float) * nr,
                     hipMemcpyDeviceToHost));
  });
}

int tc_ktimer_enable(int kernel_class, int on) {
  return guarded([&] { tc::ktimer_enable(kernel_class, on != 0); });
}

int tc_ktimer_stride(int stride) {
  return guarded([&] { tc::ktimer_set_stride(stride); });
}

int tc_graph_replays(long* count) {
  return guarded([&] {
    TC_REQUIRE(count, "graph replays: null output");
    *count = tc::graph_replays();
  });
}

int tc_alloc_stats(long* device_mallocs, long* cache_hits,
                   double* cached_bytes) {
  return guarded([&] {
    tc::arena_stats(device_mallocs, cache_hits, cached_bytes);
  });
}

int tc_release_cached(void) {
  return guarded([&] { tc::arena_trim(); });
}

int tc_ktimer_collect(int kernel_class, double* total_ms, long* count,
                      double* bytes) {
  return guarded([&] {
    double t = 0.0, b = 0.0;
    long c = 0;
    tc::ktimer_collect(kernel_class, &t, &c, &b);
    if (total_ms) *total_ms = t;
    if (count) *count = c;
    if (bytes) *bytes = b;
  });
}

}  // extern "C"
