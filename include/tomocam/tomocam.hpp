// Header-only C++ layer with the reference's tomocam:: interface
// (src/tomocam.h:44-217, src/dist_array.h, src/toeplitz.h, src/nufft.h) over
// the C-ABI of libtomocam_hip.so.
//
// Host arrays in, host arrays out (blocking), like the reference's DArray
// interface; T = float (the instantiation its Python path uses, on the tuned
// FP32 kernels) or T = double (the reference's second instantiation, on the
// library's FP64 kernels, padded widths up to 2047).  Errors throw
// std::runtime_error / std::invalid_argument instead of exit().
// Link with -ltomocam_hip and the HIP runtime (-lamdhip64).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../tomocam_hip.h"

namespace tomocam {

namespace detail {

inline void check(int status, const char* what) {
  if (status == TC_OK) return;
  std::string msg = std::string(what) + ": " + tc_last_error();
  if (status == TC_ERR_INVALID || status == TC_ERR_VALUE)
    throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// device copy of a host buffer (RAII)
template <typename T>
class Dev {
 public:
  explicit Dev(size_t n) : n_(n) {
    if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
  }
  Dev(const T* h, size_t n) : Dev(n) {
    if (n)
      hip_check(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice),
                "hipMemcpy");
  }
  ~Dev() {
    if (p_) (void)hipFree(p_);
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  T* get() const { return p_; }
  void to_host(T* h) const {
    if (n_)
      hip_check(hipMemcpy(h, p_, n_ * sizeof(T), hipMemcpyDeviceToHost),
                "hipMemcpy");
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

template <typename T>
constexpr void require_real() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "tomocam:: is instantiated for float and double");
}

template <typename T>
constexpr void require_float() {
  static_assert(std::is_same<T, float>::value,
                "this form exists for float only");
}

// The C-ABI entry point of each operation, by element type: the float
// forms take the tuned FP32 kernels, the double forms the *_f64 ones.
inline int c_project(const float* a, float* b, int S, int N, const float* ang,
                     int P) {
  return tc_project(a, b, S, N, ang, P, nullptr);
}
inline int c_project(const double* a, double* b, int S, int N,
                     const double* ang, int P) {
  return tc_project_f64(a, b, S, N, ang, P, nullptr);
}
inline int c_backproject(const float* a, float* b, int S, int P, int N,
                         const float* ang, int fbp) {
  return tc_backproject(a, b, S, P, N, ang, fbp, nullptr);
}
inline int c_backproject(const double* a, double* b, int S, int P, int N,
                         const double* ang, int fbp) {
  return tc_backproject_f64(a, b, S, P, N, ang, fbp, nullptr);
}
inline int c_gradient(const float* x, const float* b, float* g, int S, int N,
                      const float* ang, int P) {
  return tc_gradient(x, b, g, S, N, ang, P, nullptr);
}
inline int c_gradient(const double* x, const double* b, double* g, int S,
                      int N, const double* ang, int P) {
  return tc_gradient_f64(x, b, g, S, N, ang, P, nullptr);
}
inline int c_function_value(const float* x, const float* y, int S, int N,
                            const float* ang, int P, double* v) {
  return tc_function_value(x, y, S, N, ang, P, v, nullptr);
}
inline int c_function_value(const double* x, const double* y, int S, int N,
                            const double* ang, int P, double* v) {
  return tc_function_value_f64(x, y, S, N, ang, P, v, nullptr);
}
inline int c_add_total_var2(const float* x, float* g, int S, int R, int C,
                            double sigma, double p) {
  return tc_add_total_var2(x, g, S, R, C, (float)sigma, (float)p, nullptr);
}
inline int c_add_total_var2(const double* x, double* g, int S, int R, int C,
                            double sigma, double p) {
  return tc_add_total_var2_f64(x, g, S, R, C, sigma, p, nullptr);
}
inline int c_add_total_var(const float* x, float* g, int S, int R, int C,
                           double sigma, double p) {
  return tc_add_total_var(x, g, S, R, C, (float)sigma, (float)p, nullptr);
}
inline int c_add_total_var(const double* x, double* g, int S, int R, int C,
                           double sigma, double p) {
  return tc_add_total_var_f64(x, g, S, R, C, sigma, p, nullptr);
}
inline int c_preproc(const float* a, float* b, int S, int P, int N,
                     double center, int* npad) {
  return tc_preproc(a, b, S, P, N, (float)center, npad, nullptr);
}
inline int c_preproc(const double* a, double* b, int S, int P, int N,
                     double center, int* npad) {
  return tc_preproc_f64(a, b, S, P, N, center, npad, nullptr);
}
inline int c_postproc(const float* a, float* b, int S, int n, int m) {
  return tc_postproc(a, b, S, n, m, nullptr);
}
inline int c_postproc(const double* a, double* b, int S, int n, int m) {
  return tc_postproc_f64(a, b, S, n, m, nullptr);
}
inline int c_pad2d(const float* a, float* b, int S, int n, int pad) {
  return tc_pad2d(a, b, S, n, pad, nullptr);
}
inline int c_pad2d(const double* a, double* b, int S, int n, int pad) {
  return tc_pad2d_f64(a, b, S, n, pad, nullptr);
}
inline int c_xerror(const float* a, const float* b, long n, double* v) {
  return tc_xerror(a, b, n, v, nullptr);
}
inline int c_xerror(const double* a, const double* b, long n, double* v) {
  return tc_xerror_f64(a, b, n, v, nullptr);
}

// PSF handle of each element type
template <typename T>
struct Psf;
template <>
struct Psf<float> {
  using obj = tc_psf_s;
  static obj* create(const float* ang, int P, int N) {
    tc_psf h = nullptr;
    check(tc_psf_create(&h, ang, P, N, nullptr), "PointSpreadFunction");
    return h;
  }
  static void destroy(obj* h) { tc_psf_destroy(h); }
  static int gradient2(obj* h, const float* x, const float* b, float* g,
                       int S) {
    return tc_gradient2(h, x, b, g, S, nullptr);
  }
  static int function_value2(obj* h, const float* x, const float* b, int S,
                             double sq, double* v) {
    return tc_function_value2(h, x, b, S, sq, v, nullptr);
  }
};
template <>
struct Psf<double> {
  using obj = tc_psf64_s;
  static obj* create(const double* ang, int P, int N) {
    tc_psf64 h = nullptr;
    check(tc_psf_create_f64(&h, ang, P, N, nullptr), "PointSpreadFunction");
    return h;
  }
  static void destroy(obj* h) { tc_psf_destroy_f64(h); }
  static int gradient2(obj* h, const double* x, const double* b, double* g,
                       int S) {
    return tc_gradient2_f64(h, x, b, g, S, nullptr);
  }
  static int function_value2(obj* h, const double* x, const double* b, int S,
                             double sq, double* v) {
    return tc_function_value2_f64(h, x, b, S, sq, v, nullptr);
  }
};

}  // namespace detail

// src/common.h: x = slices, y = rows, z = columns
struct dim3_t {
  int x = 0, y = 0, z = 0;
  dim3_t() = default;
  dim3_t(int a, int b, int c) : x(a), y(b), z(c) {}
};

enum class PadType { LEFT, RIGHT, SYMMETRIC };

// Host volume (src/dist_array.h:41-256, the parts the hot path uses).
template <typename T>
class DArray {
 public:
  DArray() = default;
  explicit DArray(dim3_t d) : dims_(d), buf_((size_t)d.x * d.y * d.z) {}
  dim3_t dims() const { return dims_; }
  int nslices() const { return dims_.x; }
  int nrows() const { return dims_.y; }
  int ncols() const { return dims_.z; }
  size_t size() const { return buf_.size(); }
  size_t bytes() const { return buf_.size() * sizeof(T); }
  T* begin() { return buf_.data(); }
  T* end() { return buf_.data() + buf_.size(); }
  const T* begin() const { return buf_.data(); }
  const T* end() const { return buf_.data() + buf_.size(); }
  T& operator[](size_t i) { return buf_[i]; }
  const T& operator[](size_t i) const { return buf_[i]; }
  T& operator()(int i, int j, int k) {
    return buf_[((size_t)i * dims_.y + j) * dims_.z + k];
  }
  const T& operator()(int i, int j, int k) const {
    return buf_[((size_t)i * dims_.y + j) * dims_.z + k];
  }
  void init(T v) { std::fill(buf_.begin(), buf_.end(), v); }
  // sum of squares (src/dist_array.h:130-133: not a root)
  T norm() const {
    double s = 0.0;
    for (T v : buf_) s += (double)v * v;
    return (T)s;
  }
  T max() const {
    T m = (T)-1.0e20;
    for (T v : buf_) m = std::max(m, v);
    return m;
  }
  T min() const {
    T m = (T)1.0e20;
    for (T v : buf_) m = std::min(m, v);
    return m;
  }

 private:
  dim3_t dims_;
  std::vector<T> buf_;
};

namespace NUFFT {
// Polar grid of a geometry (src/nufft.h:54-192); coordinates are generated
// on the device from the angles, so this only records the geometry.
template <typename T>
class Grid {
 public:
  Grid() = default;
  Grid(int nproj, int npixel, const T* angles, int /*device id*/ = 0)
      : angles_(angles, angles + nproj), npix_(npixel) {
    detail::require_real<T>();
  }
  int nprojs() const { return (int)angles_.size(); }
  int npixels() const { return npix_; }
  int size() const { return nprojs() * npix_; }
  const std::vector<T>& angles() const { return angles_; }

 private:
  std::vector<T> angles_;
  int npix_ = 0;
};
}  // namespace NUFFT

// src/toeplitz.h:33-160
template <typename T>
class PointSpreadFunction {
  using H = detail::Psf<T>;

 public:
  PointSpreadFunction() = default;
  explicit PointSpreadFunction(const NUFFT::Grid<T>& grid) {
    detail::require_real<T>();
    h_ = std::shared_ptr<typename H::obj>(
        H::create(grid.angles().data(), grid.nprojs(), grid.npixels()),
        [](typename H::obj* p) { H::destroy(p); });
    n_ = grid.npixels();
  }
  // the device path needs no per-batch FFT plans
  void create_plans(int batch_size) { batch_ = batch_size; }
  int batch_size() const { return batch_; }
  int npixels() const { return n_; }
  typename H::obj* handle() const { return h_.get(); }

 private:
  std::shared_ptr<typename H::obj> h_;
  int n_ = 0;
  int batch_ = 0;
};

// ---- src/tomocam.h:44-217 ----

template <typename T>
DArray<T> project(DArray<T>& img, const std::vector<T>& angles) {
  detail::require_real<T>();
  const int S = img.nslices(), N = img.ncols(), P = (int)angles.size();
  detail::Dev<T> d_in(img.begin(), img.size());
  DArray<T> out(dim3_t(S, P, N));
  detail::Dev<T> d_out(out.size());
  detail::check(detail::c_project(d_in.get(), d_out.get(), S, N,
                                 angles.data(), P),
                "project");
  d_out.to_host(out.begin());
  return out;
}

template <typename T>
DArray<T> backproject(DArray<T>& sino, const std::vector<T>& angles,
                      bool fbp = false) {
  detail::require_real<T>();
  const int S = sino.nslices(), P = sino.nrows(), N = sino.ncols();
  detail::Dev<T> d_in(sino.begin(), sino.size());
  DArray<T> out(dim3_t(S, N, N));
  detail::Dev<T> d_out(out.size());
  detail::check(detail::c_backproject(d_in.get(), d_out.get(), S, P, N,
                                     angles.data(), fbp ? 1 : 0),
                "backproject");
  d_out.to_host(out.begin());
  return out;
}

template <typename T>
DArray<T> gradient(DArray<T>& x, DArray<T>& sinoT,
                   const std::vector<NUFFT::Grid<T>>& grids) {
  detail::require_real<T>();
  if (grids.empty()) throw std::invalid_argument("gradient: no grid");
  const NUFFT::Grid<T>& g = grids[0];
  const int S = x.nslices(), N = x.ncols();
  detail::Dev<T> dx(x.begin(), x.size()), db(sinoT.begin(), sinoT.size());
  DArray<T> out(x.dims());
  detail::Dev<T> dg(out.size());
  detail::check(detail::c_gradient(dx.get(), db.get(), dg.get(), S, N,
                                  g.angles().data(), g.nprojs()),
                "gradient");
  dg.to_host(out.begin());
  return out;
}

template <typename T>
DArray<T> gradient2(DArray<T>& x, DArray<T>& sinoT,
                    const std::vector<PointSpreadFunction<T>>& psfs) {
  detail::require_real<T>();
  if (psfs.empty()) throw std::invalid_argument("gradient2: no PSF");
  detail::Dev<T> dx(x.begin(), x.size()), db(sinoT.begin(), sinoT.size());
  DArray<T> out(x.dims());
  detail::Dev<T> dg(out.size());
  detail::check(detail::Psf<T>::gradient2(psfs[0].handle(), dx.get(),
                                         db.get(), dg.get(), x.nslices()),
                "gradient2");
  dg.to_host(out.begin());
  return out;
}

template <typename T>
T function_value(DArray<T>& x, DArray<T>& sino,
                 const std::vector<NUFFT::Grid<T>>& grids) {
  detail::require_real<T>();
  if (grids.empty()) throw std::invalid_argument("function_value: no grid");
  const NUFFT::Grid<T>& g = grids[0];
  detail::Dev<T> dx(x.begin(), x.size()), dy(sino.begin(), sino.size());
  double v = 0.0;
  detail::check(detail::c_function_value(dx.get(), dy.get(), x.nslices(),
                                        x.ncols(), g.angles().data(),
                                        g.nprojs(), &v),
                "function_value");
  return (T)v;
}

template <typename T>
T function_value2(DArray<T>& x, DArray<T>& sinoT,
                  const std::vector<PointSpreadFunction<T>>& psfs, T sino_sq) {
  detail::require_real<T>();
  if (psfs.empty()) throw std::invalid_argument("function_value2: no PSF");
  detail::Dev<T> dx(x.begin(), x.size()), db(sinoT.begin(), sinoT.size());
  double v = 0.0;
  detail::check(detail::Psf<T>::function_value2(psfs[0].handle(), dx.get(),
                                               db.get(), x.nslices(),
                                               (double)sino_sq, &v),
                "function_value2");
  return (T)v;
}

template <typename T>
void add_total_var2(DArray<T>& x, DArray<T>& g, T sigma, T p) {
  detail::require_real<T>();
  detail::Dev<T> dx(x.begin(), x.size()), dg(g.begin(), g.size());
  detail::check(detail::c_add_total_var2(dx.get(), dg.get(), x.nslices(),
                                        x.nrows(), x.ncols(), sigma, p),
                "add_total_var2");
  dg.to_host(g.begin());
}

template <typename T>
void add_total_var(DArray<T>& x, DArray<T>& g, float sigma, float p) {
  detail::require_real<T>();
  detail::Dev<T> dx(x.begin(), x.size()), dg(g.begin(), g.size());
  detail::check(detail::c_add_total_var(dx.get(), dg.get(), x.nslices(),
                                       x.nrows(), x.ncols(), sigma, p),
                "add_total_var");
  dg.to_host(g.begin());
}

namespace gpu {
// src/gpu/tv_hessian.cu:53-65: the reference's result never reaches the host
// array; neither does this one.
template <typename T>
void add_tv_hessian(DArray<T>& g, float sigma) {
  detail::check(tc_add_tv_hessian(nullptr, (long)g.size(), sigma, nullptr),
                "add_tv_hessian");
}
}  // namespace gpu

template <typename T>
DArray<T> preproc(DArray<T>& sino, T center) {
  detail::require_real<T>();
  const int S = sino.nslices(), P = sino.nrows(), N = sino.ncols();
  int npad = 0;
  detail::check(detail::c_preproc((const T*)nullptr, (T*)nullptr, S, P, N,
                                 center, &npad),
                "preproc");
  DArray<T> out(dim3_t(S, P, N + 2 * npad));
  detail::Dev<T> d_in(sino.begin(), sino.size()), d_out(out.size());
  detail::check(detail::c_preproc(d_in.get(), d_out.get(), S, P, N, center,
                                 &npad),
                "preproc");
  d_out.to_host(out.begin());
  return out;
}

template <typename T>
DArray<T> postproc(DArray<T>& img, int npixels) {
  detail::require_real<T>();
  DArray<T> out(dim3_t(img.nslices(), npixels, npixels));
  detail::Dev<T> d_in(img.begin(), img.size()), d_out(out.size());
  detail::check(detail::c_postproc(d_in.get(), d_out.get(), img.nslices(),
                                  img.ncols(), npixels),
                "postproc");
  d_out.to_host(out.begin());
  return out;
}

// src/padding.cpp:37-91: always symmetric, whatever `type` is
template <typename T>
DArray<T> pad2d(DArray<T>& img, int npad, PadType /*type*/) {
  detail::require_real<T>();
  const int m = img.ncols() + std::abs(npad);
  DArray<T> out(dim3_t(img.nslices(), m, m));
  detail::Dev<T> d_in(img.begin(), img.size()), d_out(out.size());
  detail::check(detail::c_pad2d(d_in.get(), d_out.get(), img.nslices(),
                               img.ncols(), npad),
                "pad2d");
  d_out.to_host(out.begin());
  return out;
}

template <typename T>
T xerror(DArray<T>& a, DArray<T>& b) {
  detail::require_real<T>();
  if (a.size() != b.size()) throw std::invalid_argument("xerror: sizes differ");
  detail::Dev<T> da(a.begin(), a.size()), db(b.begin(), b.size());
  double v = 0.0;
  detail::check(detail::c_xerror(da.get(), db.get(), (long)a.size(), &v),
                "xerror");
  return (T)v;
}

// src/mbir2.cpp:40-145
template <typename T>
DArray<T> mbir2(std::optional<DArray<T>> guess, const DArray<T>& sino,
                std::vector<T> angles, T center, int num_iters, T sigma, T tol,
                T xtol) {
  detail::require_real<T>();
  const int S = sino.nslices(), P = sino.nrows(), N = sino.ncols();
  detail::Dev<T> d_sino(sino.begin(), sino.size());
  std::unique_ptr<detail::Dev<T>> d_guess;
  if (guess.has_value())
    d_guess.reset(new detail::Dev<T>(guess->begin(), guess->size()));
  DArray<T> out(dim3_t(S, N, N));
  detail::Dev<T> d_out(out.size());
  if constexpr (std::is_same<T, double>::value)
    // the reference's run2 runs num_iters iterations whatever tol and xtol
    // are (src/optimize.h:95-151); the FP64 form does the same
    detail::check(tc_mbir2_f64(d_sino.get(),
                               d_guess ? d_guess->get() : nullptr,
                               d_out.get(), S, P, N, angles.data(), center,
                               num_iters, sigma, nullptr, nullptr),
                  "mbir2");
  else
    detail::check(tc_mbir2(d_sino.get(), d_guess ? d_guess->get() : nullptr,
                           d_out.get(), S, P, N, angles.data(), center,
                           num_iters, sigma, tol, xtol, 25, 1, nullptr,
                           nullptr),
                  "mbir2");
  d_out.to_host(out.begin());
  return out;
}

// Mixed-precision form of mbir2 (no counterpart in the reference): FP16
// gradient, FP32 spectrum, intermediates, FFTs and state, FP64 dots, T x
// refreshed in FP32 every `refresh_every` iterations.  The padded width must
// have a compile-time FFT plan (std::invalid_argument otherwise).
template <typename T>
DArray<T> mbir2_mixed(std::optional<DArray<T>> guess, const DArray<T>& sino,
                      std::vector<T> angles, T center, int num_iters, T sigma,
                      int refresh_every = 25) {
  detail::require_float<T>();
  const int S = sino.nslices(), P = sino.nrows(), N = sino.ncols();
  detail::Dev<T> d_sino(sino.begin(), sino.size());
  std::unique_ptr<detail::Dev<T>> d_guess;
  if (guess)
    d_guess.reset(new detail::Dev<T>(guess->begin(), guess->size()));
  DArray<T> out(dim3_t(S, N, N));
  detail::Dev<T> d_out(out.size());
  detail::check(tc_mbir2_mixed(d_sino.get(),
                               d_guess ? d_guess->get() : nullptr,
                               d_out.get(), S, P, N, angles.data(), center,
                               num_iters, sigma, refresh_every, 0, 0, nullptr,
                               nullptr),
                "mbir2_mixed");
  d_out.to_host(out.begin());
  return out;
}

// src/mbir.cpp:38-119 (the NUFFT normal-operator variant; x0 required)
template <typename T>
DArray<T> mbir(DArray<T>& x0, DArray<T>& sino, std::vector<T> angles,
               T center, int num_iters, T sigma, T tol, T xtol) {
  detail::require_real<T>();
  const int S = sino.nslices(), P = sino.nrows(), N = sino.ncols();
  detail::Dev<T> d_sino(sino.begin(), sino.size());
  detail::Dev<T> d_x0(x0.begin(), x0.size());
  DArray<T> out(dim3_t(S, N, N));
  detail::Dev<T> d_out(out.size());
  if constexpr (std::is_same<T, double>::value)
    detail::check(tc_mbir_f64(d_sino.get(), d_x0.get(), d_out.get(), S, P, N,
                              angles.data(), center, num_iters, sigma,
                              nullptr, nullptr),
                  "mbir");
  else
    detail::check(tc_mbir(d_sino.get(), d_x0.get(), d_out.get(), S, P, N,
                          angles.data(), center, num_iters, sigma, tol, xtol,
                          1, nullptr, nullptr),
                  "mbir");
  d_out.to_host(out.begin());
  return out;
}

}  // namespace tomocam
