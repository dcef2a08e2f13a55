// extern "C" entry points of the FP64 forms (declared in
// include/tomocam_hip.h, section "double precision").
#include <memory>

#include "fp64.h"

namespace {

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

struct tc_psf64_s {
  std::unique_ptr<tc::f64::Psf> psf;
};

static tc::f64::Psf& get_psf64(tc_psf64 p) {
  TC_REQUIRE(p && p->psf, "null FP64 PSF handle");
  return *p->psf;
}

static void fill_info64(const tc::f64::Result& r, double* h) {
  if (!h) return;
  h[0] = r.lipschitz;
  h[1] = r.step;
  h[2] = r.sino_norm;
  h[3] = r.max_sino;
  h[4] = r.padded;
  h[5] = r.iterations;
}

int tc_project_f64(const double* d_img, double* d_sino, int nslices, int npix,
                   const double* h_angles, int nproj, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_img && d_sino && h_angles, "null argument");
    tc::f64::project(d_img, d_sino, nslices, npix, h_angles, nproj,
                     as_stream(stream));
  });
}

int tc_backproject_f64(const double* d_sino, double* d_img, int nslices,
                       int nproj, int npix, const double* h_angles, int fbp,
                       void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_sino && d_img && h_angles, "null argument");
    tc::f64::backproject(d_sino, d_img, nslices, nproj, npix, h_angles,
                         fbp != 0, as_stream(stream));
  });
}

int tc_gradient_f64(const double* d_x, const double* d_sinoT, double* d_g,
                    int nslices, int npix, const double* h_angles, int nproj,
                    void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_sinoT && d_g && h_angles, "null argument");
    TC_REQUIRE(nslices >= 1 && npix >= 2 && nproj >= 1,
               "FP64 gradient: bad shape");
    tc::f64::gradient_nufft(d_x, d_sinoT, d_g, nslices, npix, h_angles, nproj,
                            as_stream(stream));
  });
}

int tc_function_value_f64(const double* d_x, const double* d_sino,
                          int nslices, int npix, const double* h_angles,
                          int nproj, double* h_out, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_sino && h_angles && h_out, "null argument");
    TC_REQUIRE(nslices >= 1 && npix >= 2 && nproj >= 1,
               "FP64 function_value: bad shape");
    hipStream_t s = as_stream(stream);
    const long n = (long)nslices * nproj * npix;
    tc::DevBuf proj(sizeof(double) * n);
    tc::f64::project(d_x, proj.as<double>(), nslices, npix, h_angles, nproj,
                     s);
    // sum (P x - y)^2 = Px.Px - 2 Px.y + y.y, in one pass each
    const double a = tc::f64::dot(proj.as<double>(), proj.as<double>(), n, s);
    const double b = tc::f64::dot(proj.as<double>(), d_sino, n, s);
    const double c = tc::f64::dot(d_sino, d_sino, n, s);
    *h_out = a - 2.0 * b + c;
  });
}

int tc_psf_create_f64(tc_psf64* out, const double* h_angles, int nproj,
                      int npix, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(out && h_angles, "null argument");
    TC_REQUIRE(nproj >= 1 && npix >= 2, "FP64 PSF: bad geometry");
    auto h = std::make_unique<tc_psf64_s>();
    h->psf = std::make_unique<tc::f64::Psf>(h_angles, nproj, npix,
                                             as_stream(stream));
    TC_HIP(hipStreamSynchronize(as_stream(stream)));
    *out = h.release();
  });
}

int tc_psf_destroy_f64(tc_psf64 p) {
  return tc::guarded([&] { delete p; });
}

int tc_toeplitz_apply_f64(tc_psf64 p, const double* d_x, double* d_out,
                          int nslices, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_out, "null argument");
    get_psf64(p).apply(d_x, nullptr, d_out, nslices, as_stream(stream));
  });
}

int tc_gradient2_f64(tc_psf64 p, const double* d_x, const double* d_sinoT,
                     double* d_g, int nslices, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_sinoT && d_g, "null argument");
    get_psf64(p).apply(d_x, d_sinoT, d_g, nslices, as_stream(stream));
  });
}

int tc_function_value2_f64(tc_psf64 p, const double* d_x,
                           const double* d_sinoT, int nslices, double sino_sq,
                           double* h_out, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_sinoT && h_out, "null argument");
    TC_REQUIRE(nslices >= 1, "FP64 function_value2: need at least one slice");
    tc::f64::Psf& P = get_psf64(p);
    hipStream_t s = as_stream(stream);
    const long n = (long)nslices * P.n() * P.n();
    tc::DevBuf tx(sizeof(double) * n);
    P.apply(d_x, nullptr, tx.as<double>(), nslices, s);
    const double xtx = tc::f64::dot(d_x, tx.as<double>(), n, s);
    const double xb = tc::f64::dot(d_x, d_sinoT, n, s);
    *h_out = sino_sq + xtx - 2.0 * xb;
  });
}

int tc_add_total_var2_f64(const double* d_x, double* d_g, int nslices,
                          int nrows, int ncols, double sigma, double p,
                          void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_g, "null argument");
    tc::f64::add_total_var(d_x, d_g, nslices, nrows, ncols, sigma, p,
                           as_stream(stream));
  });
}

int tc_add_total_var_f64(const double* d_x, double* d_g, int nslices,
                         int nrows, int ncols, double sigma, double p,
                         void* stream) {
  return tc_add_total_var2_f64(d_x, d_g, nslices, nrows, ncols, sigma, p,
                               stream);
}

int tc_qggmrf_grad_free_f64(const double* d_x, double* d_g, int nslices,
                            int nrows, int ncols, double sigma, double p,
                            void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && d_g, "null argument");
    tc::f64::add_qggmrf_grad_free(d_x, d_g, nslices, nrows, ncols, sigma, p,
                                  as_stream(stream));
  });
}

int tc_qggmrf_value_free_f64(const double* d_x, int nslices, int nrows,
                             int ncols, double sigma, double p,
                             double* h_value, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_x && h_value, "null argument");
    *h_value = tc::f64::qggmrf_value_free(d_x, nslices, nrows, ncols, sigma,
                                          p, as_stream(stream));
  });
}

int tc_mbir2_f64(const double* d_sino, const double* d_x0, double* d_recon,
                 int nslices, int nproj, int nrays, const double* h_angles,
                 double center, int num_iters, double sigma, double* h_info,
                 void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_sino && d_recon && h_angles, "null argument");
    TC_REQUIRE(sigma > 0.0, "mbir: sigma must be > 0");
    tc::f64::Result r = tc::f64::mbir2(d_sino, d_x0, d_recon, nslices, nproj,
                                       nrays, h_angles, center, num_iters,
                                       sigma, 1.2, as_stream(stream));
    fill_info64(r, h_info);
  });
}

int tc_mbir_f64(const double* d_sino, const double* d_x0, double* d_recon,
                int nslices, int nproj, int nrays, const double* h_angles,
                double center, int num_iters, double sigma, double* h_info,
                void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_sino && d_x0 && d_recon && h_angles, "null argument");
    TC_REQUIRE(sigma > 0.0, "mbir: sigma must be > 0");
    tc::f64::Result r = tc::f64::mbir_nufft(d_sino, d_x0, d_recon, nslices,
                                            nproj, nrays, h_angles, center,
                                            num_iters, sigma, 1.2,
                                            as_stream(stream));
    fill_info64(r, h_info);
  });
}

int tc_xerror_f64(const double* d_a, const double* d_b, long n, double* h_out,
                  void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_a && d_b && h_out, "null argument");
    TC_REQUIRE(n >= 1, "xerror: need at least one element");
    *h_out = tc::f64::diff_norm2(d_a, d_b, n, as_stream(stream));
  });
}

int tc_preproc_f64(const double* d_sino, double* d_out, int nslices,
                   int nproj, int nrays, double center, int* h_npad,
                   void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(nslices >= 1 && nproj >= 1 && nrays >= 1, "preproc: bad shape");
    const int npad = tc::f64::preproc(d_sino, d_out, nslices, nproj, nrays,
                                      center, as_stream(stream));
    if (h_npad) *h_npad = npad;
  });
}

int tc_postproc_f64(const double* d_img, double* d_out, int nslices, int n,
                    int npixels, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_img && d_out, "null argument");
    TC_REQUIRE(nslices >= 1, "postproc: bad shape");
    tc::f64::postproc(d_img, d_out, nslices, n, npixels, as_stream(stream));
  });
}

int tc_pad2d_f64(const double* d_img, double* d_out, int nslices, int n,
                 int padding, void* stream) {
  return tc::guarded([&] {
    TC_REQUIRE(d_img && d_out, "null argument");
    tc::f64::pad2d(d_img, d_out, nslices, n, padding, as_stream(stream));
  });
}
