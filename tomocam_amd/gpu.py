"""Device-level operators over torch CUDA tensors (float32, C-contiguous).

Thin wrappers of the C-ABI: torch supplies device memory and the current
stream; every computation runs in libtomocam_hip.so.  Mirrors the
single-device overloads of the reference (src/internals.h:44-92) and the
``PointSpreadFunction`` type (src/toeplitz.h:33-160).
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t: torch.Tensor, name: str,
         dtype: torch.dtype = torch.float32) -> C.c_void_p:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {str(dtype).replace('torch.', '')}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return C.c_void_p(t.data_ptr())


def _f64(t: torch.Tensor) -> bool:
    """float64 tensors take the library's FP64 entry points (the
    reference's T = double instantiations)."""
    return isinstance(t, torch.Tensor) and t.dtype == torch.float64


def _angles64(angles) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(angles, dtype=np.float64).ravel())


def _as3d(t: torch.Tensor) -> torch.Tensor:
    return t.unsqueeze(0) if t.dim() == 2 else t


class PointSpreadFunction:
    """Toeplitz normal operator of one polar geometry (src/toeplitz.h:33)."""

    def __init__(self, angles=None, npix: Optional[int] = None, *,
                 psf_array: Optional[torch.Tensor] = None):
        self._h = C.c_void_p()
        if psf_array is not None:
            a = psf_array.contiguous()
            n1 = a.shape[-1]
            if a.dim() != 2 or a.shape[0] != n1 or n1 % 2 == 0:
                raise ValueError("psf_array must be [2n-1, 2n-1]")
            _lib.call("tc_psf_create_from_array", C.byref(self._h),
                      _dev(a, "psf_array"), (n1 + 1) // 2, _stream())
        else:
            if angles is None or npix is None:
                raise ValueError("need angles and npix, or psf_array")
            ang = np.ascontiguousarray(np.asarray(angles, dtype=np.float32))
            _lib.call("tc_psf_create", C.byref(self._h),
                      ang.ctypes.data_as(C.c_void_p), int(ang.size),
                      int(npix), _stream())
        n = C.c_int()
        m = C.c_int()
        _lib.call("tc_psf_info", self._h, C.byref(n), C.byref(m))
        self.n = n.value
        self.fft_size = m.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                _lib.fn("tc_psf_destroy")(h)
            except Exception:
                pass
            self._h = C.c_void_p()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def array(self) -> torch.Tensor:
        n1 = 2 * self.n - 1
        out = torch.empty((n1, n1), dtype=torch.float32, device="cuda")
        _lib.call("tc_psf_get_array", self._h, _dev(out, "out"), _stream())
        return out

    def _check(self, x: torch.Tensor, name: str) -> torch.Tensor:
        x = _as3d(x)
        if x.shape[-1] != self.n or x.shape[-2] != self.n:
            raise ValueError(f"{name} must be [S, {self.n}, {self.n}]")
        return x

    def convolve(self, x: torch.Tensor, with_dot: bool = False):
        """T x (PointSpreadFunction::convolve); optionally also x . T x."""
        x = self._check(x, "x")
        out = torch.empty_like(x)
        dot = C.c_double()
        _lib.call("tc_toeplitz_apply", self._h, _dev(x, "x"),
                  _dev(out, "out"), x.shape[0],
                  C.byref(dot) if with_dot else None, _stream())
        return (out, dot.value) if with_dot else out

    def tune(self, nslices: int, mixed: bool = False) -> int:
        """Times the Toeplitz passes' slice-group sizes for ``nslices`` slices
        and keeps the fastest for the process; returns it (-1: tuned before,
        0: nothing to tune)."""
        g = C.c_int()
        _lib.call("tc_psf_tune", self._h, int(nslices), int(bool(mixed)),
                  C.byref(g), _stream())
        return g.value

    def tune_length(self, nslices: int = 1) -> int:
        """Times the eligible compile-time FFT lengths (where two plans fit
        this width) on ``nslices`` slices, rebuilds the spectrum at the
        faster one and keeps it for this width for the process; returns the
        handle's FFT length."""
        m = C.c_int()
        _lib.call("tc_psf_tune_length", self._h, int(nslices), C.byref(m),
                  _stream())
        self.fft_size = m.value
        return m.value

    def convolve_mixed(self, x: torch.Tensor, with_dot: bool = False):
        """T x with x rounded to FP16 under a per-slice power-of-two scale and
        FP16 intermediates (the mixed-precision MBIR form reads an FP16 input
        but keeps FP32 intermediates); optionally also x16 . T x16."""
        x = self._check(x, "x")
        out = torch.empty_like(x)
        dot = C.c_double()
        _lib.call("tc_toeplitz_apply_mixed", self._h, _dev(x, "x"),
                  _dev(out, "out"), x.shape[0],
                  C.byref(dot) if with_dot else None, _stream())
        return (out, dot.value) if with_dot else out

    def gradient2(self, x: torch.Tensor, sinoT: torch.Tensor) -> torch.Tensor:
        """T x - sinoT (tomocam::gradient2, src/gradient2.cpp:62)."""
        x = self._check(x, "x")
        sinoT = self._check(sinoT, "sinoT")
        g = torch.empty_like(x)
        _lib.call("tc_gradient2", self._h, _dev(x, "x"), _dev(sinoT, "sinoT"),
                  _dev(g, "g"), x.shape[0], _stream())
        return g

    def function_value2(self, x: torch.Tensor, sinoT: torch.Tensor,
                        sino_sq: float) -> float:
        """||y||^2 + sum(x . T x - 2 x . sinoT) (src/error2.cpp:60-91)."""
        x = self._check(x, "x")
        sinoT = self._check(sinoT, "sinoT")
        out = C.c_double()
        _lib.call("tc_function_value2", self._h, _dev(x, "x"),
                  _dev(sinoT, "sinoT"), x.shape[0], float(sino_sq),
                  C.byref(out), _stream())
        return out.value


def add_total_var2(x: torch.Tensor, g: torch.Tensor, sigma: float,
                   p: float = 1.2) -> torch.Tensor:
    """g += qGGMRF gradient of x, in place (tomocam::add_total_var2)."""
    x3 = _as3d(x)
    g3 = _as3d(g)
    if x3.shape != g3.shape:
        raise ValueError("x and g must have the same shape")
    S, R, Cc = x3.shape
    if _f64(x3):
        _lib.call("tc_add_total_var2_f64", _dev(x3, "x", torch.float64),
                  _dev(g3, "g", torch.float64), S, R, Cc, float(sigma),
                  float(p), _stream())
        return g
    _lib.call("tc_add_total_var2", _dev(x3, "x"), _dev(g3, "g"), S, R, Cc,
              float(sigma), float(p), _stream())
    return g


def add_total_var(x: torch.Tensor, g: torch.Tensor, sigma: float,
                  p: float = 1.2) -> torch.Tensor:
    """tomocam::add_total_var (src/tvd_update.cpp); same arithmetic."""
    x3 = _as3d(x)
    g3 = _as3d(g)
    if x3.shape != g3.shape:
        raise ValueError("x and g must have the same shape")
    S, R, Cc = x3.shape
    if _f64(x3):
        _lib.call("tc_add_total_var_f64", _dev(x3, "x", torch.float64),
                  _dev(g3, "g", torch.float64), S, R, Cc, float(sigma),
                  float(p), _stream())
        return g
    _lib.call("tc_add_total_var", _dev(x3, "x"), _dev(g3, "g"), S, R, Cc,
              float(sigma), float(p), _stream())
    return g


def qggmrf_grad_free(x: torch.Tensor, g: torch.Tensor, sigma: float,
                     p: float = 1.2) -> torch.Tensor:
    """g += exact gradient of the free-boundary qGGMRF prior, in place."""
    x3 = _as3d(x)
    g3 = _as3d(g)
    if x3.shape != g3.shape:
        raise ValueError("x and g must have the same shape")
    S, R, Cc = x3.shape
    if _f64(x3):
        _lib.call("tc_qggmrf_grad_free_f64", _dev(x3, "x", torch.float64),
                  _dev(g3, "g", torch.float64), S, R, Cc, float(sigma),
                  float(p), _stream())
        return g
    _lib.call("tc_qggmrf_grad_free", _dev(x3, "x"), _dev(g3, "g"), S, R, Cc,
              float(sigma), float(p), _stream())
    return g


def qggmrf_value_free(x: torch.Tensor, sigma: float, p: float = 1.2) -> float:
    """Free-boundary qGGMRF prior value (FP64 accumulation)."""
    x3 = _as3d(x)
    S, R, Cc = x3.shape
    v = C.c_double()
    if _f64(x3):
        _lib.call("tc_qggmrf_value_free_f64", _dev(x3, "x", torch.float64), S,
                  R, Cc, float(sigma), float(p), C.byref(v), _stream())
        return v.value
    _lib.call("tc_qggmrf_value_free", _dev(x3, "x"), S, R, Cc, float(sigma),
              float(p), C.byref(v), _stream())
    return v.value


def qggmrf_value_free_tensor(x: torch.Tensor, sigma: float,
                             p: float = 1.2) -> torch.Tensor:
    """The free-boundary prior value as a 0-d float64 device tensor,
    computed without a host synchronisation (float64 input: the FP64
    kernels, whose value comes through the host)."""
    x3 = _as3d(x)
    S, R, Cc = x3.shape
    if _f64(x3):
        return torch.tensor(qggmrf_value_free(x3, sigma, p),
                            dtype=torch.float64, device=x3.device)
    out = torch.empty((), dtype=torch.float64, device=x3.device)
    _lib.call("tc_qggmrf_value_free_dev", _dev(x3, "x"), S, R, Cc,
              float(sigma), float(p), _dev(out, "value", torch.float64),
              _stream())
    return out


def tv_hessian_zero(sigma: float) -> float:
    """Value gpu::add_tv_hessian would add (it never reaches the host)."""
    v = C.c_float()
    _lib.call("tc_add_tv_hessian", None, 0, float(sigma), C.byref(v))
    return v.value


def run2(psf: PointSpreadFunction, x0: torch.Tensor, sinoT: torch.Tensor,
         num_iters: int, step: float, sigma: float, p: float = 1.2,
         sino_sq: float = 0.0, refresh_every: int = 25,
         verbose: bool = False) -> Tuple[torch.Tensor, np.ndarray,
                                         np.ndarray, dict]:
    """Optimizer::run2 with the mbir2 objective, device resident."""
    x0 = psf._check(x0, "x0")
    sinoT = psf._check(sinoT, "sinoT")
    out = torch.empty_like(x0)
    err = np.zeros(max(num_iters, 1), dtype=np.float64)
    xerr = np.zeros(max(num_iters, 1), dtype=np.float64)
    stats = np.zeros(4, dtype=np.int32)
    _lib.call("tc_run2", psf.handle, _dev(x0, "x0"), _dev(sinoT, "sinoT"),
              _dev(out, "out"), x0.shape[0], int(num_iters), float(step),
              float(sigma), float(p), float(sino_sq), int(refresh_every),
              int(bool(verbose)),
              err.ctypes.data_as(_lib.c_double_p),
              xerr.ctypes.data_as(_lib.c_double_p),
              stats.ctypes.data_as(_lib.c_int_p), _stream())
    st = dict(zip(("iterations", "attempts", "backtracks", "toeplitz"),
                  stats.tolist()))
    return out, err[:num_iters], xerr[:num_iters], st


def graph_replays() -> int:
    """MBIR attempts this process replayed from HIP graphs."""
    c = C.c_long()
    _lib.call("tc_graph_replays", C.byref(c))
    return c.value


def ktimer_enable(kernel_class: int, on: bool = True) -> None:
    _lib.call("tc_ktimer_enable", int(kernel_class), int(bool(on)))


def ktimer_stride(stride: int = 1) -> None:
    """Time one launch in `stride` of each enabled class (timing events cost
    GPU time per launch; 1 times every launch)."""
    if int(stride) < 1:
        raise ValueError("stride must be >= 1")
    _lib.call("tc_ktimer_stride", int(stride))


KERNEL_CLASSES = {
    0: "toeplitz_rows_fwd",
    1: "toeplitz_cols",
    2: "toeplitz_rows_inv",
    3: "mbir_gradient",
    4: "mbir_commit",
    5: "qggmrf",
    6: "toeplitz_rows_fwd_f16",
    7: "toeplitz_cols_f16",
    8: "toeplitz_rows_inv_f16",
}


def ktimer_collect(kernel_class: int) -> Tuple[float, int, float]:
    """(total ms, launches, algorithmic bytes) since the last collect."""
    ms = C.c_double()
    cnt = C.c_long()
    nb = C.c_double()
    _lib.call("tc_ktimer_collect", int(kernel_class), C.byref(ms),
              C.byref(cnt), C.byref(nb))
    return ms.value, cnt.value, nb.value


def alloc_stats() -> Tuple[int, int, float]:
    """(device allocations made by the library's caches, requests served
    from a cache, bytes currently cached)."""
    m = C.c_long()
    h = C.c_long()
    b = C.c_double()
    _lib.call("tc_alloc_stats", C.byref(m), C.byref(h), C.byref(b))
    return m.value, h.value, b.value


def release_cached() -> None:
    """Frees the library's cached device memory (this thread's and idle
    caches)."""
    _lib.call("tc_release_cached")


def _angles(angles) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(angles, dtype=np.float32).ravel())


def project(img: torch.Tensor, angles) -> torch.Tensor:
    """Radon transform (tomocam::project, src/forward.cpp:33-52):
    [S, N, N] -> [S, P, N]."""
    img = _as3d(img)
    S, N, N2 = img.shape
    if N != N2:
        raise ValueError("images must be square")
    if _f64(img):
        ang = _angles64(angles)
        out = torch.empty((S, ang.size, N), dtype=torch.float64,
                          device=img.device)
        _lib.call("tc_project_f64", _dev(img, "img", torch.float64),
                  _dev(out, "sino", torch.float64), S, N,
                  ang.ctypes.data_as(C.c_void_p), int(ang.size), _stream())
        return out
    ang = _angles(angles)
    out = torch.empty((S, ang.size, N), dtype=torch.float32, device=img.device)
    _lib.call("tc_project", _dev(img, "img"), _dev(out, "sino"), S, N,
              ang.ctypes.data_as(C.c_void_p), int(ang.size), _stream())
    return out


def backproject(sino: torch.Tensor, angles, fbp: bool = False) -> torch.Tensor:
    """Adjoint of :func:`project` (tomocam::backproject,
    src/backward.cpp:42-67), Ram-Lak filtered if ``fbp``: [S, P, N] ->
    [S, N, N]."""
    sino = _as3d(sino)
    S, P, N = sino.shape
    if _f64(sino):
        ang = _angles64(angles)
        if ang.size != P:
            raise ValueError(f"got {ang.size} angles for {P} projections")
        out = torch.empty((S, N, N), dtype=torch.float64, device=sino.device)
        _lib.call("tc_backproject_f64", _dev(sino, "sino", torch.float64),
                  _dev(out, "img", torch.float64), S, P, N,
                  ang.ctypes.data_as(C.c_void_p), int(bool(fbp)), _stream())
        return out
    ang = _angles(angles)
    if ang.size != P:
        raise ValueError(f"got {ang.size} angles for {P} projections")
    out = torch.empty((S, N, N), dtype=torch.float32, device=sino.device)
    _lib.call("tc_backproject", _dev(sino, "sino"), _dev(out, "img"), S, P, N,
              ang.ctypes.data_as(C.c_void_p), int(bool(fbp)), _stream())
    return out


def gradient(x: torch.Tensor, sinoT: torch.Tensor, angles) -> torch.Tensor:
    """Re(A^H A x)/N^3 - sinoT via two NUFFTs (tomocam::gradient)."""
    x = _as3d(x)
    sinoT = _as3d(sinoT)
    S, N, _ = x.shape
    g = torch.empty_like(x)
    if _f64(x):
        ang = _angles64(angles)
        _lib.call("tc_gradient_f64", _dev(x, "x", torch.float64),
                  _dev(sinoT, "sinoT", torch.float64),
                  _dev(g, "g", torch.float64), S, N,
                  ang.ctypes.data_as(C.c_void_p), int(ang.size), _stream())
        return g
    ang = _angles(angles)
    _lib.call("tc_gradient", _dev(x, "x"), _dev(sinoT, "sinoT"), _dev(g, "g"),
              S, N, ang.ctypes.data_as(C.c_void_p), int(ang.size), _stream())
    return g


def function_value(x: torch.Tensor, sino: torch.Tensor, angles) -> float:
    """sum (P x - y)^2 (tomocam::function_value, src/error.cpp:57-59)."""
    x = _as3d(x)
    sino = _as3d(sino)
    S, N, _ = x.shape
    out = C.c_double()
    if _f64(x):
        ang = _angles64(angles)
        _lib.call("tc_function_value_f64", _dev(x, "x", torch.float64),
                  _dev(sino, "sino", torch.float64), S, N,
                  ang.ctypes.data_as(C.c_void_p), int(ang.size), C.byref(out),
                  _stream())
        return out.value
    ang = _angles(angles)
    _lib.call("tc_function_value", _dev(x, "x"), _dev(sino, "sino"), S, N,
              ang.ctypes.data_as(C.c_void_p), int(ang.size), C.byref(out),
              _stream())
    return out.value


def _kind(t: torch.Tensor):
    """(entry-point suffix, dtype) of the float32 or float64 form."""
    return ("_f64", torch.float64) if _f64(t) else ("", torch.float32)


def xerror(a: torch.Tensor, b: torch.Tensor) -> float:
    """sum (a - b)^2 (tomocam::xerror)."""
    if a.shape != b.shape:
        raise ValueError("shape mismatch")
    sfx, dt = _kind(a)
    out = C.c_double()
    _lib.call("tc_xerror" + sfx, _dev(a, "a", dt), _dev(b, "b", dt),
              a.numel(), C.byref(out), _stream())
    return out.value


def preproc(sino: torch.Tensor, center: float) -> torch.Tensor:
    """Zero-pad and centre the sinogram (tomocam::preproc)."""
    sino = _as3d(sino)
    S, P, N = sino.shape
    sfx, dt = _kind(sino)
    npad = C.c_int()
    _lib.call("tc_preproc" + sfx, None, None, S, P, N, float(center),
              C.byref(npad), _stream())
    out = torch.empty((S, P, N + 2 * npad.value), dtype=dt,
                      device=sino.device)
    _lib.call("tc_preproc" + sfx, _dev(sino, "sino", dt), _dev(out, "out", dt),
              S, P, N, float(center), C.byref(npad), _stream())
    return out


def postproc(img: torch.Tensor, npixels: int) -> torch.Tensor:
    """Crop the padded reconstruction (tomocam::postproc)."""
    img = _as3d(img)
    S, n, _ = img.shape
    sfx, dt = _kind(img)
    out = torch.empty((S, npixels, npixels), dtype=dt, device=img.device)
    _lib.call("tc_postproc" + sfx, _dev(img, "img", dt), _dev(out, "out", dt),
              S, n, int(npixels), _stream())
    return out


def pad2d(img: torch.Tensor, padding: int) -> torch.Tensor:
    """Symmetric zero padding by ``padding`` in total (tomocam::pad2d)."""
    img = _as3d(img)
    S, n, _ = img.shape
    sfx, dt = _kind(img)
    m = n + abs(int(padding))
    out = torch.empty((S, m, m), dtype=dt, device=img.device)
    _lib.call("tc_pad2d" + sfx, _dev(img, "img", dt), _dev(out, "out", dt),
              S, n, int(padding), _stream())
    return out


# Accepted iterations between FP32 recomputations of T sol, in both forms:
# the recurrence T sol += s T g drifts by FP32 rounding only (the mixed
# form's T g reads the same FP16-rounded gradient as the step, with the
# FP32 spectrum and intermediates).  At 64 x 2048^2 and 100 iterations 25
# instead of 10 is 3.5 % faster and moves the result against FP64 from
# 6.9e-7 to 7.4e-7 (FP32) and from 6.46e-6 to 6.48e-6 (mixed).
REFRESH_EVERY = 25


def mbir(sino: torch.Tensor, angles, center: float, num_iters: int = 100,
         sigma: float = 100.0, tol: float = 1e-4, xtol: float = 1e-4,
         refresh_every: Optional[int] = None, verbose: bool = False,
         guess: Optional[torch.Tensor] = None, precision: str = "fp32",
         lipschitz_iters: int = 0):
    """tomocam::mbir2 on a device-resident sinogram [S, P, N] -> [S, N, N].

    ``guess`` ([S, N, N]) replaces the FBP initial guess.  Returns (recon,
    info) with info = lipschitz, step, sino_norm, max_sino, padded width and
    accepted iterations.

    ``precision="mixed"`` runs the mixed-precision form (FP16 gradient,
    FP32 spectrum, intermediates, state and FFT arithmetic, FP64 dots, T sol
    refreshed in FP32 every ``refresh_every`` accepted iterations, default
    25 in both forms); info then also has ``saturated`` (slice-attempts whose FP16
    gradient clipped) and ``backtracks``.  ``lipschitz_iters > 0`` (mixed
    only) takes L from that many power-iteration steps on T instead of
    max(T 1).

    A float64 sinogram (or ``precision="fp64"``) runs tomocam::mbir2<double>
    (``tc_mbir2_f64``): every kernel in FP64, T sol recomputed each
    iteration, the reference's stopping rule; ``tol``/``xtol``/
    ``refresh_every`` do not apply there."""
    if precision not in ("fp32", "mixed", "fp64"):
        raise ValueError("precision must be 'fp32', 'mixed' or 'fp64'")
    if _f64(sino) or precision == "fp64":
        # the FP64 instantiation (tomocam::mbir2<double>): float64 in and out
        if precision == "mixed" or lipschitz_iters:
            raise ValueError("float64 input runs the FP64 form")
        return _mbir_f64("tc_mbir2_f64", sino, angles, center, num_iters,
                         sigma, guess)
    if lipschitz_iters and precision != "mixed":
        raise ValueError("lipschitz_iters is a mixed-precision option")
    if refresh_every is None:
        refresh_every = REFRESH_EVERY
    sino = _as3d(sino)
    S, P, N = sino.shape
    ang = _angles(angles)
    if ang.size != P:
        raise ValueError(f"got {ang.size} angles for {P} projections")
    x0 = None
    if guess is not None:
        guess = _as3d(guess)
        if tuple(guess.shape) != (S, N, N):
            raise ValueError(f"guess must be [{S}, {N}, {N}]")
        x0 = _dev(guess, "guess")
    out = torch.empty((S, N, N), dtype=torch.float32, device=sino.device)
    if precision == "mixed":
        info = np.zeros(8, dtype=np.float64)
        _lib.call("tc_mbir2_mixed", _dev(sino, "sino"), x0, _dev(out, "recon"),
                  S, P, N, ang.ctypes.data_as(C.c_void_p), float(center),
                  int(num_iters), float(sigma), int(refresh_every),
                  int(lipschitz_iters), int(bool(verbose)),
                  info.ctypes.data_as(_lib.c_double_p), _stream())
        keys = ("lipschitz", "step", "sino_norm", "max_sino", "padded",
                "iterations", "saturated", "backtracks")
        return out, dict(zip(keys, info.tolist()))
    info = np.zeros(6, dtype=np.float64)
    _lib.call("tc_mbir2", _dev(sino, "sino"), x0, _dev(out, "recon"), S, P, N,
              ang.ctypes.data_as(C.c_void_p), float(center), int(num_iters),
              float(sigma), float(tol), float(xtol), int(refresh_every),
              int(bool(verbose)), info.ctypes.data_as(_lib.c_double_p),
              _stream())
    keys = ("lipschitz", "step", "sino_norm", "max_sino", "padded",
            "iterations")
    return out, dict(zip(keys, info.tolist()))


def mbir_nufft(sino: torch.Tensor, angles, center: float,
               x0: torch.Tensor, num_iters: int = 100, sigma: float = 100.0,
               tol: float = 1e-4, xtol: float = 1e-4,
               verbose: bool = False):
    """tomocam::mbir (src/mbir.cpp:38-119), the NUFFT normal-operator MBIR
    without a PSF: sinogram [S, P, N] and initial guess x0 [S, N, N] (both
    device) -> (recon [S, N, N], info)."""
    sino = _as3d(sino)
    if _f64(sino):
        return _mbir_f64("tc_mbir_f64", sino, angles, center, num_iters,
                         sigma, x0)
    S, P, N = sino.shape
    ang = _angles(angles)
    if ang.size != P:
        raise ValueError(f"got {ang.size} angles for {P} projections")
    x0 = _as3d(x0)
    if tuple(x0.shape) != (S, N, N):
        raise ValueError(f"x0 must be [{S}, {N}, {N}]")
    out = torch.empty((S, N, N), dtype=torch.float32, device=sino.device)
    info = np.zeros(6, dtype=np.float64)
    _lib.call("tc_mbir", _dev(sino, "sino"), _dev(x0, "x0"), _dev(out, "recon"),
              S, P, N, ang.ctypes.data_as(C.c_void_p), float(center),
              int(num_iters), float(sigma), float(tol), float(xtol),
              int(bool(verbose)), info.ctypes.data_as(_lib.c_double_p),
              _stream())
    keys = ("lipschitz", "step", "sino_norm", "max_sino", "padded",
            "iterations")
    return out, dict(zip(keys, info.tolist()))


def _mbir_f64(entry, sino, angles, center, num_iters, sigma, x0):
    """tc_mbir2_f64 / tc_mbir_f64 on float64 device tensors."""
    sino = _as3d(sino)
    if sino.dtype != torch.float64:
        sino = sino.double().contiguous()
    S, P, N = sino.shape
    ang = _angles64(angles)
    if ang.size != P:
        raise ValueError(f"got {ang.size} angles for {P} projections")
    xp = None
    if x0 is not None:
        x0 = _as3d(x0)
        if tuple(x0.shape) != (S, N, N):
            raise ValueError(f"initial guess must be [{S}, {N}, {N}]")
        x0 = x0.double().contiguous()
        xp = _dev(x0, "x0", torch.float64)
    out = torch.empty((S, N, N), dtype=torch.float64, device=sino.device)
    info = np.zeros(6, dtype=np.float64)
    _lib.call(entry, _dev(sino, "sino", torch.float64), xp,
              _dev(out, "recon", torch.float64), S, P, N,
              ang.ctypes.data_as(C.c_void_p), float(center), int(num_iters),
              float(sigma), info.ctypes.data_as(_lib.c_double_p), _stream())
    keys = ("lipschitz", "step", "sino_norm", "max_sino", "padded",
            "iterations")
    return out, dict(zip(keys, info.tolist()))


class PointSpreadFunction64:
    """PointSpreadFunction<double> (src/toeplitz.h:33-160): the FP64 Toeplitz
    operator of one polar geometry, on float64 tensors."""

    def __init__(self, angles, npix: int):
        self._h = C.c_void_p()
        ang = _angles64(angles)
        _lib.call("tc_psf_create_f64", C.byref(self._h),
                  ang.ctypes.data_as(C.c_void_p), int(ang.size), int(npix),
                  _stream())
        self.n = int(npix)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                _lib.fn("tc_psf_destroy_f64")(h)
            except Exception:
                pass
            self._h = C.c_void_p()

    def _check(self, x, name):
        x = _as3d(x)
        if x.shape[-1] != self.n or x.shape[-2] != self.n:
            raise ValueError(f"{name} must be [S, {self.n}, {self.n}]")
        return x

    def convolve(self, x: torch.Tensor) -> torch.Tensor:
        x = self._check(x, "x")
        out = torch.empty_like(x)
        _lib.call("tc_toeplitz_apply_f64", self._h, _dev(x, "x", torch.float64),
                  _dev(out, "out", torch.float64), x.shape[0], _stream())
        return out

    def gradient2(self, x: torch.Tensor, sinoT: torch.Tensor) -> torch.Tensor:
        x = self._check(x, "x")
        sinoT = self._check(sinoT, "sinoT")
        g = torch.empty_like(x)
        _lib.call("tc_gradient2_f64", self._h, _dev(x, "x", torch.float64),
                  _dev(sinoT, "sinoT", torch.float64),
                  _dev(g, "g", torch.float64), x.shape[0], _stream())
        return g

    def function_value2(self, x: torch.Tensor, sinoT: torch.Tensor,
                        sino_sq: float) -> float:
        x = self._check(x, "x")
        sinoT = self._check(sinoT, "sinoT")
        v = C.c_double()
        _lib.call("tc_function_value2_f64", self._h,
                  _dev(x, "x", torch.float64),
                  _dev(sinoT, "sinoT", torch.float64), x.shape[0],
                  float(sino_sq), C.byref(v), _stream())
        return v.value


def mbir_batch(stacks, num_iters: int = 100, sigma: float = 100.0,
               refresh_every: int = REFRESH_EVERY):
    """Uneven batching: reconstruct many stacks in one library call.

    ``stacks`` is a sequence of (sino, angles, center) with sino a CUDA
    tensor [S, P, N] (or [P, N]); sizes, angle sets and slice counts may all
    differ.  Each stack is exactly :func:`mbir` with the FBP initial guess.
    Returns (list of recon tensors [S, N, N], list of info dicts)."""
    n = len(stacks)
    sinos, recons, angs = [], [], []
    ns = np.zeros(max(n, 1), dtype=np.int32)
    nps = np.zeros_like(ns)
    nrs = np.zeros_like(ns)
    cens = np.zeros(max(n, 1), dtype=np.float32)
    for k, (sino, angles, center) in enumerate(stacks):
        s3 = _as3d(sino)
        S, P, N = s3.shape
        a = _angles(angles)
        if a.size != P:
            raise ValueError(f"stack {k}: {a.size} angles for {P} projections")
        sinos.append(s3)
        angs.append(a)
        recons.append(torch.empty((S, N, N), dtype=torch.float32,
                                  device=s3.device))
        ns[k], nps[k], nrs[k], cens[k] = S, P, N, float(center)
    sp = (C.c_void_p * max(n, 1))(*[_dev(t, "sino").value for t in sinos])
    rp = (C.c_void_p * max(n, 1))(*[_dev(t, "recon").value for t in recons])
    ap = (C.c_void_p * max(n, 1))(*[a.ctypes.data for a in angs])
    info = np.zeros(6 * max(n, 1), dtype=np.float64)
    _lib.call("tc_mbir2_batch", n, C.cast(sp, C.c_void_p),
              C.cast(rp, C.c_void_p),
              ns.ctypes.data_as(C.POINTER(C.c_int)),
              nps.ctypes.data_as(C.POINTER(C.c_int)),
              nrs.ctypes.data_as(C.POINTER(C.c_int)),
              C.cast(ap, C.c_void_p),
              cens.ctypes.data_as(C.c_void_p), int(num_iters), float(sigma),
              int(refresh_every),
              info.ctypes.data_as(C.POINTER(C.c_double)), _stream())
    keys = ("lipschitz", "step", "sino_norm", "max_sino", "padded",
            "iterations")
    infos = [dict(zip(keys, info[6 * k:6 * k + 6].tolist())) for k in range(n)]
    return recons, infos
