"""FP32 CPU restatement of ``tomocam::mbir2`` used as the benchmark's CPU
baseline.

CPU-baseline infrastructure only (imported by ``bench.py``'s cpu_baseline
leg and by tests); the product never imports it.  The reference has no CPU path (it
links only cufinufft and cuFFT), so this is a project-written restatement of
its algorithm, built to be fast on the host rather than a strawman:

* the algorithm of ``src/mbir2.cpp:40-145`` and ``Optimizer::run2``
  (``src/optimize.h:96-154``) in FP32, with four Toeplitz applications per
  attempt exactly as the reference does (gradient, error(sol), error(y),
  error(sol) after acceptance);
* ES-kernel NUFFT (w = 8, beta = 2.30 w, upsampling 2: FINUFFT's parameters
  for the reference's 1.2e-7 tolerance) with C/OpenMP spreading and
  interpolation (``oracle/cpu_kernels.c``) and scipy.fft (pocketfft) on all
  cores;
* Toeplitz products by real FFTs of a smooth size >= 2n-1 (scipy.fft, all
  cores);
* the qGGMRF stencil of ``test/tvd.cpp:39-75`` in C/OpenMP.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import time
from typing import Optional

import numpy as np
import scipy.fft as sfft

from . import tomocam_oracle as O

W = 8
BETA = 2.30 * W
_HERE = os.path.dirname(os.path.abspath(__file__))
LIBNAME = "libtc_cpu.so"
_lib = None


def workers() -> int:
    return os.cpu_count() or 1


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, LIBNAME)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.qggmrf_add_f32.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int,
                                     C.c_float, C.c_float]
        L.es_spread_f32.argtypes = [vp, vp, vp, C.c_long, C.c_int, C.c_int,
                                    C.c_int, C.c_float, vp]
        L.es_interp_f32.argtypes = [vp, vp, vp, C.c_long, C.c_int, C.c_int,
                                    C.c_int, C.c_float, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


class Geometry:
    """Polar grid in fine-grid units (u, v) for ``nmodes`` modes."""

    def __init__(self, angles, npix: int, nmodes: int):
        self.npix = npix
        self.nmodes = nmodes
        self.nf = O.smooth_size(max(2 * nmodes, 2 * W))
        x, y = O.polar_grid(np.asarray(angles, dtype=np.float64), npix)
        s = self.nf / (2 * np.pi)
        self.u = np.ascontiguousarray(x.ravel() * s + self.nf / 2)
        self.v = np.ascontiguousarray(y.ravel() * s + self.nf / 2)
        k = O.modes(nmodes)
        corr = O.ESKernel(W, BETA).correction(k, self.nf)
        self.corr = (corr * np.where(k % 2 == 0, 1.0, -1.0)).astype(np.float32)
        self.kidx = k % self.nf


def nufft1_real(c: np.ndarray, g: Geometry, scale: float) -> np.ndarray:
    """Re(type 1) * scale; c is complex64 [S, J]."""
    S, J = c.shape
    nf = g.nf
    fine = np.empty((S, nf, nf), dtype=np.complex64)
    cc = np.ascontiguousarray(c, dtype=np.complex64)
    lib().es_spread_f32(_p(cc), _p(g.u), _p(g.v), J, S, nf, W, BETA, _p(fine))
    fh = sfft.ifft2(fine, axes=(-2, -1), norm="forward", workers=workers())
    sub = fh[:, g.kidx][:, :, g.kidx].real
    return (sub * (g.corr[:, None] * g.corr[None, :]) * scale).astype(np.float32)


def nufft2_real(f: np.ndarray, g: Geometry) -> np.ndarray:
    """Type 2 of real f [S, n, n] -> complex64 [S, J]."""
    S = f.shape[0]
    nf = g.nf
    grid = np.zeros((S, nf, nf), dtype=np.complex64)
    grid[:, g.kidx[:, None], g.kidx[None, :]] = (
        f * (g.corr[:, None] * g.corr[None, :]))
    b = np.ascontiguousarray(sfft.fft2(grid, axes=(-2, -1), workers=workers()))
    J = g.u.size
    out = np.empty((S, J), dtype=np.complex64)
    lib().es_interp_f32(_p(b), _p(g.u), _p(g.v), J, S, nf, W, BETA, _p(out))
    return out


def _ray_dft(v: np.ndarray, sign: int) -> np.ndarray:
    u = sfft.ifftshift(v, axes=-1)
    if sign < 0:
        u = sfft.fft(u, axis=-1, workers=workers())
    else:
        u = sfft.ifft(u, axis=-1, norm="forward", workers=workers())
    return sfft.fftshift(u, axes=-1)


def backproject(sino: np.ndarray, angles, fbp: bool = False) -> np.ndarray:
    S, P, n = sino.shape
    v = _ray_dft(sino.astype(np.complex64), -1)
    if fbp:
        v = v * O.ramlak(n).astype(np.float32)[None, None, :]
    g = Geometry(angles, n, n)
    return nufft1_real(v.reshape(S, P * n), g, 1.0 / (n * n))


class Toeplitz:
    def __init__(self, psf_arr: np.ndarray, n: int):
        self.n = n
        self.M = O.smooth_size(2 * n - 1)
        circ = np.zeros((self.M, self.M), dtype=np.float32)
        idx = np.arange(-(n - 1), n) % self.M
        circ[np.ix_(idx, idx)] = psf_arr
        self.H = sfft.rfft2(circ, workers=workers()).real.astype(np.float32) / (
            float(n) ** 3)

    def __call__(self, x: np.ndarray) -> np.ndarray:
        M = self.M
        X = sfft.rfft2(x, s=(M, M), workers=workers())
        X *= self.H[None]
        y = sfft.irfft2(X, s=(M, M), workers=workers())
        return np.ascontiguousarray(y[:, :self.n, :self.n], dtype=np.float32)


def qggmrf_add(x: np.ndarray, g: np.ndarray, sigma: float, p: float) -> None:
    x = np.ascontiguousarray(x, dtype=np.float32)
    S, R, Cc = x.shape
    lib().qggmrf_add_f32(_p(x), _p(g), S, R, Cc, float(sigma), float(p))


class Mbir2CPU:
    """tomocam::mbir2 restated in FP32 on the CPU (reference control flow),
    split into the set-up (constructor) and resumable iterations."""

    def __init__(self, sino: np.ndarray, angles, center: float,
                 sigma: float, p: float = 1.2):
        t0 = time.perf_counter()
        sino = np.asarray(sino, dtype=np.float32)
        maxv = sino.max()
        if maxv == 0:
            raise ValueError("sinogram has ZERO max value")
        sino2 = sino / maxv
        x0 = backproject(sino2, angles, fbp=True)
        self.nrays = sino.shape[-1]
        sino2 = O.preproc(sino2, center).astype(np.float32)
        n2 = sino2.shape[-1]
        x0 = O.pad2d(x0, n2 - self.nrays).astype(np.float32)
        self.sinoT = backproject(sino2, angles)
        self.sino_norm = np.float32(np.sum(sino2.astype(np.float64) ** 2))
        t1 = time.perf_counter()
        # the PSF, its spectrum and L: once per geometry (stack), not per
        # slice
        g1 = Geometry(angles, n2, 2 * n2 - 1)
        ones = np.ones((1, g1.u.size), dtype=np.complex64)
        psf_arr = nufft1_real(ones, g1, 1.0)[0]
        self.T = Toeplitz(psf_arr, n2)
        L = float(self.T(np.ones((1, n2, n2), dtype=np.float32)).max())
        self.step0 = np.float32(min(1.0, 1.0 / L))
        t2 = time.perf_counter()
        self.sigma, self.p = sigma, p
        self.sol = x0
        self.x = x0
        self.t = 1.0
        self.tnew = 1.0
        self.step = self.step0
        self.timing = dict(setup_slices_s=t1 - t0, setup_psf_s=t2 - t1)

    def _grad(self, x):
        gg = self.T(x) - self.sinoT
        qggmrf_add(x, gg, self.sigma, self.p)
        return gg

    def _err(self, x):
        tx = self.T(x)
        return math.sqrt(float(self.sino_norm
                               + np.sum(x * tx, dtype=np.float32)
                               - 2 * np.sum(x * self.sinoT, dtype=np.float32)))

    def iterate(self, num_iters: int) -> None:
        """``num_iters`` accepted iterations of Optimizer::run2
        (src/optimize.h:96-154), four Toeplitz products per attempt."""
        sol, x, t, tnew, step = self.sol, self.x, self.t, self.tnew, self.step
        step0 = self.step0
        for _ in range(num_iters):
            while True:
                beta = np.float32(tnew * (1 / t - 1))
                tnew = 0.5 * (math.sqrt(t ** 4 + 4 * t ** 2) - t ** 2)
                y = sol + (sol - x) * beta
                gr = self._grad(y)
                sol = y - gr * step
                fx = self._err(sol)
                fy = self._err(y)
                gy = 0.5 * step * float(np.sum(gr * gr, dtype=np.float32))
                if fx > fy + gy:
                    step = np.float32(step * 0.9)
                else:
                    step = step0
                    t = tnew
                    x = sol
                    break
            self._err(sol)
        self.sol, self.x, self.t, self.tnew, self.step = sol, x, t, tnew, step

    def result(self) -> np.ndarray:
        return O.postproc(self.sol, self.nrays)


def mbir2(sino: np.ndarray, angles, center: float, num_iters: int,
          sigma: float, p: float = 1.2) -> np.ndarray:
    """tomocam::mbir2 (src/mbir2.cpp:40-145) in FP32 on the CPU."""
    m = Mbir2CPU(sino, angles, center, sigma, p)
    m.iterate(num_iters)
    return m.result()


def sample_rate(sino: np.ndarray, angles, center: float, num_iters: int,
                sigma: float, sample_slices: int, sample_iters: int,
                repeats: int = 3, total_slices: Optional[int] = None) -> dict:
    """Throughput of the CPU restatement on a bounded sample: the set-up of
    ``sample_slices`` slices once, one warm-up iteration, then ``repeats``
    timed runs of ``sample_iters`` iterations, whose median is taken.
    Extrapolated linearly to ``num_iters`` iterations per slice of the whole
    stack (slices are independent units of work; the PSF is once per
    stack)."""
    total = total_slices or sino.shape[0]
    sample_slices = max(1, min(sample_slices, sino.shape[0]))
    sub = np.ascontiguousarray(sino[:sample_slices])
    m = Mbir2CPU(sub, angles, center, sigma)
    m.iterate(1)  # warm-up (FFT plans, thread pools)
    runs = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        m.iterate(sample_iters)
        runs.append(time.perf_counter() - t0)
    it_s = float(np.median(runs)) / sample_iters  # per iteration, sample
    tm = m.timing
    per_slice = (tm["setup_slices_s"] / sample_slices
                 + tm["setup_psf_s"] / total
                 + it_s / sample_slices * num_iters)
    return dict(slices_per_s=1.0 / per_slice, setup_s=tm["setup_slices_s"],
                psf_s=tm["setup_psf_s"], iter_s=it_s, runs_s=runs,
                sample_slices=sample_slices, sample_iters=sample_iters,
                threads=workers())
