/*
 * C-ABI of the MI355X-native tomoCAM MBIR hot path (libtomocam_hip.so).
 *
 * Conventions
 *   - Every function returns an int status (TC_OK = 0); on failure
 *     tc_last_error() returns a thread-local message.  Nothing calls exit().
 *   - Pointers named d_* are device pointers on the current HIP device and
 *     h_* are host pointers.  Arrays are C-contiguous float32:
 *       image stacks  [nslices][npix][npix]
 *       sinograms     [nslices][nproj][nrays]
 *   - `stream` is a hipStream_t (NULL = the default stream).  Functions that
 *     return a host scalar, and the MBIR drivers, synchronise `stream`;
 *     others are asynchronous on `stream`.
 *   - The library never frees caller memory.  Opaque handles (tc_psf) are
 *     created and destroyed explicitly.
 *
 * Each declaration names the reference interface it replaces
 * (file:line in lbl-camera/tomocam).
 */
#ifndef TOMOCAM_HIP_H
#define TOMOCAM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_OK 0
#define TC_ERR_INVALID 1  /* bad argument or shape */
#define TC_ERR_HIP 2      /* HIP runtime error */
#define TC_ERR_NOMEM 3    /* device allocation failed */
#define TC_ERR_INTERNAL 4 /* anything else */
#define TC_ERR_VALUE 5    /* numerically invalid input (all-zero sinogram) */

typedef struct tc_psf_s* tc_psf;
typedef struct tc_psf64_s* tc_psf64;

/* Message of the last failure on this thread ("" if none). */
const char* tc_last_error(void);

/* Library version. */
int tc_version(int* major, int* minor, int* patch);

/* cTomocam.set_num_of_gpus (src/pyTomocam.cpp:141-142, src/machine.h:48-54).
 * The library runs on one GPU: n must be >= 1 and is otherwise ignored. */
int tc_set_num_of_gpus(int n);

/* ---------------------------------------------------------------------- */
/* projectors                                                               */
/* ---------------------------------------------------------------------- */

/* tomocam::project (src/radon.cpp:70-96, single device src/forward.cpp:33-52):
 * d_sino[S][P][N] = Re(fftshift(IFFT(ifftshift(A x))))/N^2 with A the type-2
 * NUFFT on the polar grid r_j = 2 pi (j - N/2)/(N - 1) of h_angles[P]. */
int tc_project(const float* d_img, float* d_sino, int nslices, int npix,
               const float* h_angles, int nproj, void* stream);

/* tomocam::backproject (src/iradon.cpp:75-102, single device
 * src/backward.cpp:42-67): exact adjoint of tc_project; with fbp != 0 the
 * Ram-Lak ramp |k - N/2|/N is applied first (src/gpu/ramlak.cu:14-31).
 * d_sino[S][P][N] -> d_img[S][N][N]. */
int tc_backproject(const float* d_sino, float* d_img, int nslices, int nproj,
                   int npix, const float* h_angles, int fbp, void* stream);

/* tomocam::gradient (src/gradient.cpp:35-103): d_g = Re(A^H A x)/N^3 - d_sinoT
 * by a type-2 and a type-1 NUFFT (the PSF-free normal operator). */
int tc_gradient(const float* d_x, const float* d_sinoT, float* d_g,
                int nslices, int npix, const float* h_angles, int nproj,
                void* stream);

/* tomocam::function_value (src/error.cpp:36-88): *h_out = sum (P x - y)^2
 * (FP64 accumulation). */
int tc_function_value(const float* d_x, const float* d_sino, int nslices,
                      int npix, const float* h_angles, int nproj,
                      double* h_out, void* stream);

/* tomocam::xerror (src/xerror.cpp:37-85): *h_out = sum (a - b)^2. */
int tc_xerror(const float* d_a, const float* d_b, long n, double* h_out,
              void* stream);

/* tomocam::preproc (src/preproc.cpp:73-108): zero-pad each sinogram row by
 * npad on both sides and roll by the centre offset.  *h_npad receives npad
 * (output width N + 2 npad); call with d_out == NULL to query it. */
int tc_preproc(const float* d_sino, float* d_out, int nslices, int nproj,
               int nrays, float center, int* h_npad, void* stream);

/* tomocam::postproc (src/postproc.cpp:69-102): crop [S][n][n] to the centred
 * [S][npixels][npixels]. */
int tc_postproc(const float* d_img, float* d_out, int nslices, int n,
                int npixels, void* stream);

/* tomocam::pad2d (src/padding.cpp:37-91): zero-pad [S][n][n] by `padding`
 * in total, symmetrically, to [S][n+padding][n+padding]. */
int tc_pad2d(const float* d_img, float* d_out, int nslices, int n,
             int padding, void* stream);

/* ---------------------------------------------------------------------- */
/* Toeplitz (point-spread-function) normal operator                         */
/* ---------------------------------------------------------------------- */

/* PointSpreadFunction<T>(grid) (src/toeplitz.h:51-73): builds the PSF of the
 * polar geometry (h_angles[nproj], npix rays per angle) with a type-1 NUFFT
 * on (2 npix - 1)^2 modes, and its real spectrum. */
int tc_psf_create(tc_psf* out, const float* h_angles, int nproj, int npix,
                  void* stream);

/* Same object from an explicit real PSF array d_psf[2n-1][2n-1] (PSF(m) at
 * index m + n - 1), e.g. one computed elsewhere. */
int tc_psf_create_from_array(tc_psf* out, const float* d_psf, int n,
                             void* stream);

/* Copies the real PSF array [2n-1][2n-1] of `p` into d_psf (device). */
int tc_psf_get_array(tc_psf p, float* d_psf, void* stream);

/* ~PointSpreadFunction (src/toeplitz.h:75-81). */
int tc_psf_destroy(tc_psf p);

/* Image width n and FFT size M used by the convolution. */
int tc_psf_info(tc_psf p, int* n, int* fft_size);

/* The FFT size M a PSF of width n would use (M >= 2n - 1; the reference
 * uses 2n - 1, src/toeplitz.h:57-60): the smallest length with compile-time
 * passes when within 10 % of the smallest even 2,3,5,7-smooth length, else
 * that.  No GPU needed. */
int tc_toeplitz_fft_length(int n, int* fft_size);

/* PointSpreadFunction::convolve (src/toeplitz.h:95-101):
 * d_out = T d_x = sum_q x[q] PSF(a - q) / n^3 per slice.  If h_xTx is not
 * NULL it receives x . T x (FP64, computed in the spectral domain) and the
 * call synchronises `stream`. */
int tc_toeplitz_apply(tc_psf p, const float* d_x, float* d_out, int nslices,
                      double* h_xTx, void* stream);

/* tc_toeplitz_apply without allocating (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace (the row-transformed intermediate, the dot partials and their
 * sum) lives in a 72 MiB array that the library reserves statically on
 * each GPU for the life of the process; slices are processed as many at a
 * time as fit.  A call needs one slice's intermediate, 8 (M/2 + 1) n' bytes
 * rounded up (n' the width rounded up to even; 67.9 MB at n = 2908), plus
 * 8 bytes per dot partial when h_xTx is requested; larger requests fail
 * with TC_ERR_INVALID before any work.  The array is shared by every call
 * on a GPU and keeps the last call's contents: calls that overlap on one
 * GPU (concurrent streams or threads) race.  With h_xTx non-NULL the call
 * synchronises `stream`. */
int tc_toeplitz_apply_noalloc(tc_psf p, const float* d_x, float* d_out,
                              int nslices, double* h_xTx, void* stream);

/* Autotuner of the Toeplitz passes (no counterpart in the reference): times
 * one application on scratch data (FP32 or the mixed-precision passes) for
 * slice-group sizes of once, half and twice the default (a group's
 * intermediate stays in the memory-side cache between the three passes; the
 * default keeps it near 144 MB) on one stream, and for once and half the
 * default with groups alternating over a second stream (one group's row
 * passes overlap the next group's column pass).  Candidates are taken in
 * that order and a later one must be at least 2 % faster to replace the
 * best, so near-ties resolve the same way on every run.  The choice holds
 * for this width, FFT length and slice count for the rest of the process;
 * TOMOCAM_HIP_TOEPLITZ_STREAMS=1 or 2 overrides the stream count.  *group
 * (optional) receives the group size, -1 if the shape was tuned before, 0
 * if there is nothing to tune (no compile-time plan, or one slice).
 * tc_mbir2 / tc_mbir2_mixed tune on first use unless the environment sets
 * TOMOCAM_HIP_AUTOTUNE=0.  Synchronises `stream`. */
int tc_psf_tune(tc_psf p, int nslices, int mixed, int* group, void* stream);

/* Autotuner of the FFT length, and so of the compile-time radix plan (no
 * counterpart in the reference): where two compile-time plans fit the
 * width (within 10 % of the smallest smooth length >= 2n - 1, e.g. 750 and
 * 768 for n = 350 ... 362, 2250 and 2304, 4500 and 4608), times one FP32
 * application of each on `nslices` slices of scratch data and rebuilds the
 * handle's spectrum at the faster one; the static rule's choice is kept
 * unless the other is at least 2 % faster.  The choice holds for this width
 * for the rest of the process (later handles and tc_toeplitz_fft_length
 * use it).  *fft_size (optional) receives the handle's FFT length.
 * tc_mbir2 / tc_mbir2_mixed run it on first use of a width unless
 * TOMOCAM_HIP_AUTOTUNE=0.  Synchronises `stream`. */
int tc_psf_tune_length(tc_psf p, int nslices, int* fft_size, void* stream);

/* Mixed-precision T x (no counterpart in the reference): x rounded to FP16
 * under a per-slice power-of-two scale, the three passes computing in FP32
 * with FP16 intermediates and FP16 spectrum (the FP16-storage operator; the
 * mixed MBIR form applies T to its FP16 gradient with FP32 intermediates
 * and spectrum).  Needs an FFT length with compile-time
 * passes.  Optional h_xTx as for tc_toeplitz_apply, here for the rounded x.
 * Synchronises `stream`. */
int tc_toeplitz_apply_mixed(tc_psf p, const float* d_x, float* d_out,
                            int nslices, double* h_xTx, void* stream);

/* tomocam::gradient2 (src/gradient2.cpp:36-100, src/tomocam.h:85-86):
 * d_g = T d_x - d_sinoT. */
int tc_gradient2(tc_psf p, const float* d_x, const float* d_sinoT, float* d_g,
                 int nslices, void* stream);

/* tomocam::function_value2 (src/error2.cpp:40-94, src/tomocam.h:109-110):
 * *h_out = sino_sq + sum (x . T x - 2 x . sinoT), FP64 accumulation. */
int tc_function_value2(tc_psf p, const float* d_x, const float* d_sinoT,
                       int nslices, double sino_sq, double* h_out,
                       void* stream);

/* ---------------------------------------------------------------------- */
/* qGGMRF regulariser                                                       */
/* ---------------------------------------------------------------------- */

/* tomocam::add_total_var2 (src/tv_update.cpp:75-106) with qGGMRF_kernel
 * (src/gpu/cuTV2.cu:32-93): d_g += sum_d w_d rho'(x_v - x_clamp(v+d)),
 * q = 2, c = 0.001, replicate-at-edge clamp on all three axes. */
int tc_add_total_var2(const float* d_x, float* d_g, int nslices, int nrows,
                      int ncols, float sigma, float p, void* stream);

/* tomocam::add_total_var (src/tvd_update.cpp:82-117, tvd_update_kernel
 * src/gpu/cuTVD.cu:32-190): the same arithmetic as add_total_var2. */
int tc_add_total_var(const float* d_x, float* d_g, int nslices, int nrows,
                     int ncols, float sigma, float p, void* stream);

/* gpu::add_tv_hessian (src/gpu/tv_hessian.cu:53-65).  The reference adds
 * sum_d w_d q / (sigma^q c) to a device copy that is discarded, so the host
 * array is unchanged; this mirrors that (h_g is not modified) and returns
 * the value it would have added in *h_value (may be NULL). */
int tc_add_tv_hessian(float* h_g, long n, float sigma, float* h_value);

/* Differentiable form (no counterpart in the reference; it serves the
 * torch.autograd projector/prior of tomocam_amd.autograd).  Free-boundary
 * qGGMRF prior R(x) = sum over unordered in-bounds neighbour pairs
 * {v, v+d} of w_d rho(x_v - x_{v+d}), rho(d) = |d/s|^2 / (c + |d/s|^(2-p)),
 * with the reference's weights and constants:
 *   tc_qggmrf_grad_free:  d_g += dR/dx (exact gradient of that sum; in the
 *                         interior it equals tc_add_total_var2);
 *   tc_qggmrf_value_free: *h_value = R(x), accumulated in FP64.
 * Synchronises `stream` (value only). */
int tc_qggmrf_grad_free(const float* d_x, float* d_g, int nslices, int nrows,
                        int ncols, float sigma, float p, void* stream);
int tc_qggmrf_value_free(const float* d_x, int nslices, int nrows, int ncols,
                         float sigma, float p, double* h_value, void* stream);
/* The same value into the device double *d_value, asynchronously on
 * `stream` (for callers that keep the loss on the device). */
int tc_qggmrf_value_free_dev(const float* d_x, int nslices, int nrows,
                             int ncols, float sigma, float p, double* d_value,
                             void* stream);

/* ---------------------------------------------------------------------- */
/* MBIR iteration                                                           */
/* ---------------------------------------------------------------------- */

/* Optimizer::run2 (src/optimize.h:96-154) with the mbir2 objective
 * (src/mbir2.cpp:122-136): accelerated gradient with backtracking on
 * 0.5 ||P x - y||^2 + qGGMRF, device resident.  d_x0, d_sinoT, d_out are
 * [nslices][n][n] with n = PSF width; sino_sq = ||y||^2.  tol/xtol are
 * accepted and ignored like the reference.  refresh_every = k recomputes
 * T(sol) and T(x) every k accepted steps (0 = never).  Optional outputs:
 * h_err[num_iters], h_xerr[num_iters] (the reference's printed error and
 * x-err per iteration); h_stats[4] = {accepted, attempts, backtracks,
 * Toeplitz applications}.  Synchronises `stream`. */
int tc_run2(tc_psf p, const float* d_x0, const float* d_sinoT, float* d_out,
            int nslices, int num_iters, double step, float sigma, float p_exp,
            double sino_sq, int refresh_every, int verbose, double* h_err,
            double* h_xerr, int* h_stats, void* stream);

/* tomocam::mbir2 (src/mbir2.cpp:40-145), as called by cTomocam.mbir
 * (src/pyTomocam.cpp:116-134): normalise by max(sino), initial guess (the
 * FBP of the normalised sinogram, or d_x0[S][N][N] if not NULL), pad/centre,
 * P^T y, PSF, step = min(1, 1/L), run2, crop.  d_sino[S][P][N] ->
 * d_recon[S][N][N], all on the device.  tol/xtol are ignored like the
 * reference.  An all-zero sinogram returns TC_ERR_VALUE without a guess and
 * copies the guess to d_recon with one (src/mbir2.cpp:50-53).  Optional
 * h_info[6] = {lipschitz, step, ||y||^2, max(sino), padded width, accepted
 * iterations}.  Synchronises `stream`. */
int tc_mbir2(const float* d_sino, const float* d_x0, float* d_recon,
             int nslices, int nproj, int nrays, const float* h_angles,
             float center, int num_iters, float sigma, float tol, float xtol,
             int refresh_every, int verbose, double* h_info, void* stream);

/* Mixed-precision form of tc_mbir2 (no counterpart in the reference, whose
 * Python path is FP32 throughout): the same pipeline and iteration, with
 * the gradient stored in FP16 under a per-slice power-of-two scale, the
 * PSF spectrum, FFT butterflies, Toeplitz intermediates and NUFFT
 * spreading in FP32, the iterate, T x and b in FP32, line-search dots in
 * FP64, and T sol recomputed in FP32 every refresh_every accepted
 * iterations (0 = never).  The step and T g use the same rounded gradient,
 * so the T y recurrence drifts only by FP32 rounding.
 * lipschitz_iters > 0 takes L from that many power-iteration steps on T
 * (FP64 dots) instead of max(T 1).  Needs an FFT length with compile-time
 * passes (TC_ERR_INVALID otherwise).  Optional h_info[8] = tc_mbir2's six
 * values, then the number of slice-attempts whose FP16 gradient saturated
 * and the number of backtracks. */
int tc_mbir2_mixed(const float* d_sino, const float* d_x0, float* d_recon,
                   int nslices, int nproj, int nrays, const float* h_angles,
                   float center, int num_iters, float sigma,
                   int refresh_every, int lipschitz_iters, int verbose,
                   double* h_info, void* stream);

/* tomocam::mbir (src/mbir.cpp:38-119): the NUFFT normal-operator MBIR,
 * without a PSF.  Normalises by (sino - min) / (max - min), pads and
 * centres, b = P^T y, step = min(1, 1 / max(Re(A^H A 1)/n^3)), then the
 * reference's run2 with gradient Re(A^H A y)/n^3 - b + R'(y) and error
 * ||P v - y||.  d_x0 [nslices][nrays][nrays] (device) is required, as in
 * the reference; a sinogram with max == min returns d_x0 unchanged.
 * h_info as for tc_mbir2 (padded width 0 for the unchanged return). */
int tc_mbir(const float* d_sino, const float* d_x0, float* d_recon,
            int nslices, int nproj, int nrays, const float* h_angles,
            float center, int num_iters, float sigma, float tol, float xtol,
            int verbose, double* h_info, void* stream);

/* Real-data front end, device side: projections as read from the HDF5 file
 * (tc_h5_read_projections in tomocam_h5.h), d_proj [nproj][nslices][ncols],
 * to sinograms d_sino [nslices][nproj][ncols - drop_last_col].  Replaces
 * the host transpose of tomocam::h5::Reader::read_sinogram
 * (src/hdf5/reader.h:110-114) and DArray::dropcol (src/dist_array.h:
 * 258-282), which test/mbir2.cpp:104-107 applies to even widths (the last
 * column is dropped; the centre is kept). */
int tc_projections_to_sinogram(const float* d_proj, float* d_sino, int nproj,
                               int nslices, int ncols, int drop_last_col,
                               void* stream);

/* Uneven batching (no counterpart in the reference, which runs one stack
 * per call): nstacks independent MBIR reconstructions of different image
 * sizes, angle sets and slice counts in one call, each exactly tc_mbir2
 * with an FBP initial guess.  Stack k: d_sinos[k] [nslices[k]][nproj[k]]
 * [nrays[k]] -> d_recons[k] [nslices[k]][nrays[k]][nrays[k]], angles
 * h_angles[k] (nproj[k] floats, host), centre centers[k].  Stacks run on
 * TOMOCAM_HIP_BATCH_STREAMS (default 4) streams of their own at once, one
 * host thread each, largest first (1: in order on `stream`); the call waits
 * for `stream` first and returns when every stack is done.  NUFFT plans are
 * cached by geometry.  h_info (optional): 6 doubles per stack as for
 * tc_mbir2.  On error the message names the lowest failing stack index and
 * stacks not yet started are not run. */
int tc_mbir2_batch(int nstacks, const float* const* d_sinos,
                   float* const* d_recons, const int* nslices,
                   const int* nproj, const int* nrays,
                   const float* const* h_angles, const float* centers,
                   int num_iters, float sigma, int refresh_every,
                   double* h_info, void* stream);

/* Same with host arrays (copies in and out, like the reference's DArray
 * interface). */
int tc_mbir2_host(const float* h_sino, float* h_recon, int nslices, int nproj,
                  int nrays, const float* h_angles, float center,
                  int num_iters, float sigma, float tol, float xtol,
                  int verbose);

/* ---------------------------------------------------------------------- */
/* double precision: the reference's T = double instantiations              */
/* (src/mbir2.cpp:158-167, src/mbir.cpp:124, src/radon.cpp:100,             */
/* src/iradon.cpp:107, src/gradient2.cpp:107).  Arrays are float64 and every */
/* kernel computes in FP64; the NUFFT kernel has width 16 (FINUFFT's width   */
/* for the tolerance 1e-15 the reference plans double transforms with,     */
/* src/nufft.h:219).  The qGGMRF derivative is evaluated in FP64 (the       */
/* reference's d_pot_func is float-only, src/gpu/potential_function.cuh:    */
/* 76-99, so its double regulariser carries FP32 rounding).  Widths up to   */
/* 2047 (padded), where every FFT line fits one workgroup's LDS; larger     */
/* widths fail with TC_ERR_INVALID.  Angles are float64 host arrays.        */
/* ---------------------------------------------------------------------- */

/* tomocam::project<double> (src/radon.cpp:70-100). */
int tc_project_f64(const double* d_img, double* d_sino, int nslices, int npix,
                   const double* h_angles, int nproj, void* stream);

/* tomocam::backproject<double> (src/iradon.cpp:75-107). */
int tc_backproject_f64(const double* d_sino, double* d_img, int nslices,
                       int nproj, int npix, const double* h_angles, int fbp,
                       void* stream);

/* tomocam::gradient<double> (src/gradient.cpp:35-103). */
int tc_gradient_f64(const double* d_x, const double* d_sinoT, double* d_g,
                    int nslices, int npix, const double* h_angles, int nproj,
                    void* stream);

/* tomocam::function_value<double> (src/error.cpp:36-88); synchronises. */
int tc_function_value_f64(const double* d_x, const double* d_sino,
                          int nslices, int npix, const double* h_angles,
                          int nproj, double* h_out, void* stream);

/* PointSpreadFunction<double> (src/toeplitz.h:33-160) from the geometry;
 * synchronises `stream`. */
int tc_psf_create_f64(tc_psf64* out, const double* h_angles, int nproj,
                      int npix, void* stream);
int tc_psf_destroy_f64(tc_psf64 p);

/* PointSpreadFunction<double>::convolve (src/toeplitz.h:95-101). */
int tc_toeplitz_apply_f64(tc_psf64 p, const double* d_x, double* d_out,
                          int nslices, void* stream);

/* tomocam::gradient2<double> (src/gradient2.cpp:36-107). */
int tc_gradient2_f64(tc_psf64 p, const double* d_x, const double* d_sinoT,
                     double* d_g, int nslices, void* stream);

/* tomocam::function_value2<double> (src/error2.cpp:40-94); synchronises. */
int tc_function_value2_f64(tc_psf64 p, const double* d_x,
                           const double* d_sinoT, int nslices, double sino_sq,
                           double* h_out, void* stream);

/* tomocam::add_total_var2<double> / add_total_var<double>
 * (src/tv_update.cpp:41-106, src/tvd_update.cpp:41-117). */
int tc_add_total_var2_f64(const double* d_x, double* d_g, int nslices,
                          int nrows, int ncols, double sigma, double p,
                          void* stream);
int tc_add_total_var_f64(const double* d_x, double* d_g, int nslices,
                         int nrows, int ncols, double sigma, double p,
                         void* stream);

/* FP64 forms of tc_qggmrf_grad_free / tc_qggmrf_value_free (the
 * differentiable prior; no counterpart in the reference). */
int tc_qggmrf_grad_free_f64(const double* d_x, double* d_g, int nslices,
                            int nrows, int ncols, double sigma, double p,
                            void* stream);
int tc_qggmrf_value_free_f64(const double* d_x, int nslices, int nrows,
                             int ncols, double sigma, double p,
                             double* h_value, void* stream);

/* tomocam::mbir2<double> (src/mbir2.cpp:40-145, :158-167): the Toeplitz
 * MBIR with the FBP guess when d_x0 is NULL, p = 1.2.  Optional
 * h_info[6] as tc_mbir2.  Synchronises. */
int tc_mbir2_f64(const double* d_sino, const double* d_x0, double* d_recon,
                 int nslices, int nproj, int nrays, const double* h_angles,
                 double center, int num_iters, double sigma, double* h_info,
                 void* stream);

/* tomocam::mbir<double> (src/mbir.cpp:38-124): the NUFFT normal-operator
 * MBIR; d_x0 required.  Optional h_info[6].  Synchronises. */
int tc_mbir_f64(const double* d_sino, const double* d_x0, double* d_recon,
                int nslices, int nproj, int nrays, const double* h_angles,
                double center, int num_iters, double sigma, double* h_info,
                void* stream);

/* tomocam::xerror<double> (src/xerror.cpp:65-85): sum (a - b)^2;
 * synchronises. */
int tc_xerror_f64(const double* d_a, const double* d_b, long n, double* h_out,
                  void* stream);

/* tomocam::preproc<double> (src/preproc.cpp:73-108), as tc_preproc. */
int tc_preproc_f64(const double* d_sino, double* d_out, int nslices,
                   int nproj, int nrays, double center, int* h_npad,
                   void* stream);

/* tomocam::postproc<double> (src/postproc.cpp:69-102), as tc_postproc. */
int tc_postproc_f64(const double* d_img, double* d_out, int nslices, int n,
                    int npixels, void* stream);

/* tomocam::pad2d<double> (src/padding.cpp:37-91), as tc_pad2d. */
int tc_pad2d_f64(const double* d_img, double* d_out, int nslices, int n,
                 int padding, void* stream);

/* ---------------------------------------------------------------------- */
/* instrumentation                                                          */
/* ---------------------------------------------------------------------- */

/* Kernel classes: 0 Toeplitz rows (forward), 1 Toeplitz columns,
 * 2 Toeplitz rows (inverse), 3 MBIR gradient (stencil + assembly),
 * 4 MBIR commit, 5 standalone qGGMRF, 6-8 the FP16-storage Toeplitz passes
 * of the mixed-precision form (rows forward, columns, rows inverse).  When enabled, launches of that class
 * are bracketed by HIP events on their stream. */
int tc_ktimer_enable(int kernel_class, int on);

/* Time one launch in `stride` (>= 1, default 1) of each enabled class; the
 * others run unbracketed.  Each timing event costs about 1 us of GPU time,
 * which is several per cent of a loop of 60 us launches; a stride coprime
 * to the launches per application (3 for power-of-two slice groups) samples
 * every launch position in turn.  Resets the per-class launch counters. */
int tc_ktimer_stride(int stride);

/* Sum of the event-timed durations (ms), number of launches and algorithmic
 * HBM bytes (each input read once, each output written once) of a class
 * since the previous collect. */
int tc_ktimer_collect(int kernel_class, double* total_ms, long* count,
                      double* bytes);

/* Number of MBIR attempts this process has replayed from HIP graphs (no
 * counterpart in the reference).  With TOMOCAM_HIP_GRAPHS=1 the MBIR entry
 * points capture the steady-state attempt (y = sol, full step) once per
 * buffer role and replay it, except with kernel timing enabled, an FFT
 * length without compile-time passes, the separate commit kernel, a second
 * Toeplitz stream, or concurrent tc_mbir2_batch workers.  Off by default:
 * it measured no faster than launching the kernels.  While one thread
 * captures, a synchronous HIP call on the null stream from another thread
 * of the process fails, so turn it on only where MBIR calls are the
 * process's only concurrent HIP work. */
int tc_graph_replays(long* count);

/* Device-memory arenas (no counterpart in the reference, whose every call
 * allocates and frees its temporaries).  The MBIR drivers take their
 * buffers from a per-host-thread cache, so a repeated call allocates
 * nothing and never waits in hipFree.  Device allocations made for the
 * caches so far, requests served from a cache, and the bytes cached. */
int tc_alloc_stats(long* device_mallocs, long* cache_hits,
                   double* cached_bytes);

/* Frees the cached device memory of the calling thread and of every cache
 * no thread is using. */
int tc_release_cached(void);

#ifdef __cplusplus
}
#endif

#endif /* TOMOCAM_HIP_H */
