// Exercises the tomocam:: C++ layer (include/tomocam/tomocam.hpp) the way
// the reference's test drivers do (NPRandom inputs, test/timer.h:40-56) and
// writes the results as raw float32 (and, for the double instantiations,
// float64) files into argv[1] for the Python test to compare with the
// oracle.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "tomocam/tomocam.hpp"

namespace {

// test/timer.h:40-56: mt19937(5489), 53-bit doubles from two draws
struct NPRandom {
  std::mt19937 gen{5489u};
  double rand() {
    long a = gen() >> 5;
    long b = gen() >> 6;
    return (a * 67108864.0 + b) / 9007199254740992.0;
  }
};

template <typename T>
void dump(const std::string& dir, const char* name, const T* p, size_t n) {
  std::string path = dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(p, sizeof(T), n, f);
  std::fclose(f);
}

// The T = double instantiations on a fresh NPRandom(5489) stream: outputs as
// raw float64 files (*.f64).
void double_forms(const std::string& dir) {
  NPRandom rng;
  tomocam::DArray<double> a(tomocam::dim3_t(2, 33, 40));
  for (size_t i = 0; i < a.size(); ++i) a[i] = rng.rand();
  tomocam::DArray<double> g(a.dims());
  g.init(0.0);
  tomocam::add_total_var2(a, g, 10.0, 1.2);
  dump(dir, "tv_out.f64", g.begin(), g.size());

  const int N = 32, P = 24;
  std::vector<double> ang(P);
  for (int i = 0; i < P; ++i) ang[i] = i * M_PI / P;
  tomocam::DArray<double> img(tomocam::dim3_t(2, N, N));
  for (size_t i = 0; i < img.size(); ++i) img[i] = rng.rand();
  auto sino = tomocam::project(img, ang);
  auto bp = tomocam::backproject(sino, ang, false);
  dump(dir, "sino.f64", sino.begin(), sino.size());
  dump(dir, "bp.f64", bp.begin(), bp.size());

  tomocam::NUFFT::Grid<double> grid(P, N, ang.data(), 0);
  std::vector<tomocam::PointSpreadFunction<double>> psfs;
  psfs.emplace_back(grid);
  tomocam::DArray<double> x2(img.dims());
  for (size_t i = 0; i < x2.size(); ++i) x2[i] = rng.rand();
  auto g2 = tomocam::gradient2(x2, bp, psfs);
  dump(dir, "grad2.f64", g2.begin(), g2.size());
  const double fv2 = tomocam::function_value2(x2, bp, psfs, sino.norm());
  dump(dir, "fv2.f64", &fv2, 1);
  const double xe = tomocam::xerror(img, x2);
  dump(dir, "xerror.f64", &xe, 1);

  auto pre = tomocam::preproc(sino, 13.0);
  dump(dir, "preproc.f64", pre.begin(), pre.size());
  auto pad = tomocam::pad2d(img, 10, tomocam::PadType::SYMMETRIC);
  auto back = tomocam::postproc(pad, N);
  double d = 0.0;
  for (size_t i = 0; i < img.size(); ++i) d += std::fabs(back[i] - img[i]);
  std::printf("double pad/crop round trip: %g\n", d);
  if (d != 0.0) throw std::runtime_error("pad2d/postproc round trip");

  auto rec = tomocam::mbir2<double>(std::nullopt, sino, ang, N / 2.0, 10,
                                    10.0, 0.0, 0.0);
  dump(dir, "mbir.f64", rec.begin(), rec.size());
  tomocam::DArray<double> x0(img.dims());
  x0.init(0.0);
  auto rec1 = tomocam::mbir<double>(x0, sino, ang, N / 2.0, 5, 10.0, 0.0, 0.0);
  dump(dir, "mbir_nufft.f64", rec1.begin(), rec1.size());
  std::printf("double forms ok\n");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: api_check <outdir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  try {
    NPRandom rng;
    // qGGMRF (test/tvd.cpp:80-110 shape, reduced)
    tomocam::DArray<float> a(tomocam::dim3_t(2, 33, 40));
    for (size_t i = 0; i < a.size(); ++i) a[i] = (float)rng.rand();
    tomocam::DArray<float> g(a.dims());
    g.init(0.f);
    tomocam::add_total_var2(a, g, 10.0f, 1.2f);
    dump(dir, "tv_in.f32", a.begin(), a.size());
    dump(dir, "tv_out.f32", g.begin(), g.size());

    // projectors
    const int N = 32, P = 24;
    std::vector<float> ang(P);
    for (int i = 0; i < P; ++i) ang[i] = (float)(i * M_PI / P);
    tomocam::DArray<float> img(tomocam::dim3_t(2, N, N));
    for (size_t i = 0; i < img.size(); ++i) img[i] = (float)rng.rand();
    auto sino = tomocam::project(img, ang);
    auto bp = tomocam::backproject(sino, ang, false);
    dump(dir, "img.f32", img.begin(), img.size());
    dump(dir, "sino.f32", sino.begin(), sino.size());
    dump(dir, "bp.f32", bp.begin(), bp.size());

    // Toeplitz gradient and function value
    tomocam::NUFFT::Grid<float> grid(P, N, ang.data(), 0);
    std::vector<tomocam::PointSpreadFunction<float>> psfs;
    psfs.emplace_back(grid);
    psfs[0].create_plans(4);
    // at a second image (at img itself T img - bp cancels to ~1e-3)
    tomocam::DArray<float> x2(img.dims());
    for (size_t i = 0; i < x2.size(); ++i) x2[i] = (float)rng.rand();
    dump(dir, "x2.f32", x2.begin(), x2.size());
    auto g2 = tomocam::gradient2(x2, bp, psfs);
    dump(dir, "grad2.f32", g2.begin(), g2.size());
    const float fv2 = tomocam::function_value2(x2, bp, psfs, sino.norm());
    dump(dir, "fv2.f32", &fv2, 1);

    // mbir2 without a guess (the Python path)
    auto rec = tomocam::mbir2<float>(std::nullopt, sino, ang, N / 2.0f, 10,
                                     10.0f, 1e-4f, 1e-4f);
    dump(dir, "mbir.f32", rec.begin(), rec.size());

    // the mixed-precision form needs a compile-time FFT length: width 32
    // pads to 44 (FFT length 90), which has none
    bool mixed_threw = false;
    try {
      tomocam::mbir2_mixed<float>(std::nullopt, sino, ang, N / 2.0f, 2, 10.0f);
    } catch (const std::invalid_argument&) {
      mixed_threw = true;
    }
    std::printf("mixed without plan threw: %d\n", mixed_threw ? 1 : 0);
    if (!mixed_threw) return 1;
    // and agrees with mbir2 at width 256 (padded 362, FFT length 750)
    {
      const int N2 = 256, P2 = 90;
      tomocam::DArray<float> img2(tomocam::dim3_t(1, N2, N2));
      for (int r = 0; r < N2; ++r)
        for (int c = 0; c < N2; ++c) {
          const float dx = (c - N2 / 2) / (float)N2, dy = (r - N2 / 2) / (float)N2;
          img2[(size_t)r * N2 + c] = dx * dx + 2 * dy * dy < 0.1f ? 1.f : 0.f;
        }
      std::vector<float> ang2(P2);
      for (int p = 0; p < P2; ++p) ang2[p] = (float)(p * M_PI / P2);
      auto s2 = tomocam::project(img2, ang2);
      auto r32 = tomocam::mbir2<float>(std::nullopt, s2, ang2, N2 / 2.0f, 20,
                                       100.0f, 0.f, 0.f);
      auto r16 = tomocam::mbir2_mixed<float>(std::nullopt, s2, ang2,
                                             N2 / 2.0f, 20, 100.0f);
      double num = 0.0, den = 0.0;
      for (size_t i = 0; i < r32.size(); ++i) {
        num += ((double)r16[i] - r32[i]) * ((double)r16[i] - r32[i]);
        den += (double)r32[i] * r32[i];
      }
      const double rel = std::sqrt(num / den);
      std::printf("mixed vs fp32 relL2 %.3e\n", rel);
      if (!(rel < 1e-4)) return 1;
    }

    double_forms(dir);

    // an all-zero sinogram throws (no guess)
    tomocam::DArray<float> zero(sino.dims());
    zero.init(0.f);
    bool threw = false;
    try {
      tomocam::mbir2<float>(std::nullopt, zero, ang, N / 2.0f, 2, 10.0f, 0.f,
                            0.f);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    std::printf("zero-sinogram threw: %d\n", threw ? 1 : 0);
    std::printf("api_check ok\n");
    return threw ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "api_check failed: %s\n", e.what());
    return 1;
  }
}
