"""FP64 CPU restatement of tomoCAM's MBIR hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``tomocam_amd``) imports,
links or executes this module; only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Every function restates one piece of the reference algorithm and cites the
reference file:line it follows (paths relative to the reference checkout).
The reference's own arithmetic lives in cufinufft / cuFFT, which cannot run
here, so the NUFFT is restated from its definition:

* ``nudft_type1`` / ``nudft_type2`` are the exact sums (separable matrix
  products), usable for small sizes.
* ``nufft_type1`` / ``nufft_type2`` are an exponential-of-semicircle (ES)
  kernel NUFFT (Barnett, Magland, af Klinteberg, SIAM J. Sci. Comput. 41
  (2019) C479) evaluated in FP64 with a wide kernel, checked against the exact
  sums in the tests.

Parity pins (each checked in ``tests/test_oracle.py``):

* polar grid: closed form and brute force;
* NUFFT: exact NUDFT, point-at-origin and on-grid known answers, adjointness;
* projectors: adjoint identity <Px, s> = <x, P^T s>; numpy FFT conventions;
* Toeplitz: T x == Re(A^H A x)/N^3 (brute force) for several FFT sizes;
* qGGMRF: the CPU stencil of ``test/tvd.cpp:14-75`` restated loop-for-loop on
  small volumes, finite differences of the pair potential, closed-form values;
* run2/mbir2: control flow restated line by line, invariants (accepted steps
  satisfy the Armijo-type test, exactly ``num_iters`` accepted steps).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

# ---------------------------------------------------------------------------
# random numbers of the reference test drivers
# ---------------------------------------------------------------------------


class NPRandom:
    """``test/timer.h:40-56``: mt19937(5489), 53-bit doubles from two draws.

    Identical to ``numpy.random.RandomState(5489).random_sample`` (the same
    generator and the same 27+26-bit construction); pinned in the tests by
    reimplementing the two-draw construction on raw 32-bit outputs.
    """

    def __init__(self, seed: int = 5489):
        self._rs = np.random.RandomState(seed)

    def rand(self, n: int) -> np.ndarray:
        return self._rs.random_sample(n)


def nprandom_from_words(words: np.ndarray) -> np.ndarray:
    """``test/timer.h:50-55``: ((w0 >> 5) * 2^26 + (w1 >> 6)) / 2^53."""
    w = np.asarray(words, dtype=np.uint64)
    a = (w[0::2] >> np.uint64(5)).astype(np.float64)
    b = (w[1::2] >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b) / 9007199254740992.0


# ---------------------------------------------------------------------------
# A.1 polar grid
# ---------------------------------------------------------------------------


def polar_grid(angles: np.ndarray, npix: int):
    """``src/gpu/nufft_utils.cu:37-42``.

    r_j = 2*pi*(j - floor(N/2))/(N-1); x = r cos(theta), y = r sin(theta);
    flattened p*N + j (here returned as [P, N] arrays).
    """
    c = npix // 2
    r = (2.0 * np.pi / (npix - 1)) * (np.arange(npix, dtype=np.float64) - c)
    th = np.asarray(angles, dtype=np.float64)
    return np.outer(np.cos(th), r), np.outer(np.sin(th), r)


def modes(n: int) -> np.ndarray:
    """CMCL mode order used by cufinufft: k = index - floor(n/2)."""
    return np.arange(n) - n // 2


# ---------------------------------------------------------------------------
# A.2 exact non-uniform DFTs
# ---------------------------------------------------------------------------


def nudft_type1(c: np.ndarray, x: np.ndarray, y: np.ndarray, nmodes: int,
                chunk: int = 8192) -> np.ndarray:
    """``src/nufft.h:230-308`` (type 1, iflag=+1), exact sum.

    F[k2][k1] = sum_j c_j exp(+i (k1 x_j + k2 y_j)); c is [S, J] (or [J]).
    Returns [S, nmodes, nmodes] complex128.
    """
    c = np.atleast_2d(np.asarray(c, dtype=np.complex128))
    S = c.shape[0]
    x = np.ravel(x)
    y = np.ravel(y)
    k = modes(nmodes).astype(np.float64)
    out = np.zeros((S, nmodes, nmodes), dtype=np.complex128)
    for j0 in range(0, x.size, chunk):
        xs = x[j0:j0 + chunk]
        ys = y[j0:j0 + chunk]
        e1 = np.exp(1j * np.outer(xs, k))  # [J, k1]
        e2 = np.exp(1j * np.outer(ys, k))  # [J, k2]
        for s in range(S):
            out[s] += e2.T @ (c[s, j0:j0 + chunk, None] * e1)
    return out


def nudft_type2(f: np.ndarray, x: np.ndarray, y: np.ndarray,
                chunk: int = 8192) -> np.ndarray:
    """``src/nufft.h:311-384`` (type 2, iflag=-1), exact sum.

    c_j = sum_{k1,k2} F[k2][k1] exp(-i (k1 x_j + k2 y_j)); f is [S, n, n].
    Returns [S, J] complex128.
    """
    f = np.asarray(f, dtype=np.complex128)
    if f.ndim == 2:
        f = f[None]
    S, n, _ = f.shape
    x = np.ravel(x)
    y = np.ravel(y)
    k = modes(n).astype(np.float64)
    out = np.zeros((S, x.size), dtype=np.complex128)
    for j0 in range(0, x.size, chunk):
        xs = x[j0:j0 + chunk]
        ys = y[j0:j0 + chunk]
        e1 = np.exp(-1j * np.outer(xs, k))
        e2 = np.exp(-1j * np.outer(ys, k))
        for s in range(S):
            out[s, j0:j0 + chunk] = np.sum((e2 @ f[s]) * e1, axis=1)
    return out


# ---------------------------------------------------------------------------
# A.2 ES-kernel NUFFT (FP64)
# ---------------------------------------------------------------------------


def smooth_size(n: int, even: bool = True, primes=(2, 3, 5, 7)) -> int:
    """Smallest integer >= n whose prime factors are all in ``primes``."""
    m = max(int(n), 1)
    while True:
        if (not even) or m % 2 == 0:
            r = m
            for p in primes:
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
        m += 1


@dataclass
class ESKernel:
    """phi(z) = exp(beta (sqrt(1 - z^2) - 1)) on |z| <= 1, width w grid points."""
    w: int
    beta: float

    @staticmethod
    def for_width(w: int) -> "ESKernel":
        # beta = 2.30 w for upsampling 2 (FINUFFT paper, Sec. 4)
        return ESKernel(w, 2.30 * w)

    def phi(self, z: np.ndarray) -> np.ndarray:
        z = np.asarray(z, dtype=np.float64)
        out = np.zeros_like(z)
        m = np.abs(z) <= 1.0
        out[m] = np.exp(self.beta * (np.sqrt(1.0 - z[m] ** 2) - 1.0))
        return out

    def correction(self, k: np.ndarray, nf: int) -> np.ndarray:
        """p(k) = h / psi_hat(k); psi(t) = phi(2 t / (w h)), h = 2 pi / nf.

        psi_hat(k) = (w h / 2) int_{-1}^{1} phi(z) cos(k z w h / 2) dz,
        evaluated by Gauss-Legendre quadrature.
        """
        h = 2.0 * np.pi / nf
        zq, wq = np.polynomial.legendre.leggauss(200)
        k = np.asarray(k, dtype=np.float64)
        integ = (self.phi(zq)[None, :] *
                 np.cos(np.outer(k, zq) * (self.w * h / 2.0))) @ wq
        return h / ((self.w * h / 2.0) * integ)


def _spread_weights(u: np.ndarray, ker: ESKernel):
    """Integer start index and w weights of each coordinate (grid units)."""
    w = ker.w
    l0 = np.ceil(u - w / 2.0).astype(np.int64)
    offs = np.arange(w)
    z = (l0[:, None] + offs[None, :] - u[:, None]) * (2.0 / w)
    return l0, ker.phi(z)


def nufft_type1(c: np.ndarray, x: np.ndarray, y: np.ndarray, nmodes: int,
                w: int = 14) -> np.ndarray:
    """Type 1 (iflag=+1) via spreading onto a centred fine grid of nf points.

    Fine-grid point l sits at X_l = (l - nf/2) h; F_k = p(k1) p(k2)
    (-1)^(k1+k2) IFFT_unnormalised(b)[k mod nf].
    """
    ker = ESKernel.for_width(w)
    c = np.atleast_2d(np.asarray(c, dtype=np.complex128))
    S = c.shape[0]
    nf = smooth_size(max(2 * nmodes, 2 * w))
    x = np.ravel(x)
    y = np.ravel(y)
    ux = x * (nf / (2 * np.pi)) + nf / 2
    uy = y * (nf / (2 * np.pi)) + nf / 2
    lx, wx = _spread_weights(ux, ker)
    ly, wy = _spread_weights(uy, ker)
    k = modes(nmodes)
    p = ker.correction(k, nf) * np.where(k % 2 == 0, 1.0, -1.0)
    out = np.empty((S, nmodes, nmodes), dtype=np.complex128)
    for s in range(S):
        grid = np.zeros(nf * nf, dtype=np.complex128)
        for a in range(w):
            iy = (ly + a) % nf
            for b in range(w):
                ix = (lx + b) % nf
                val = c[s] * (wy[:, a] * wx[:, b])
                idx = iy * nf + ix
                grid += (np.bincount(idx, weights=val.real, minlength=nf * nf)
                         + 1j * np.bincount(idx, weights=val.imag,
                                            minlength=nf * nf))
        fh = np.fft.ifft2(grid.reshape(nf, nf)) * (nf * nf)
        sub = fh[np.ix_(k % nf, k % nf)]
        out[s] = sub * p[:, None] * p[None, :]
    return out


def nufft_type2(f: np.ndarray, x: np.ndarray, y: np.ndarray,
                w: int = 14) -> np.ndarray:
    """Type 2 (iflag=-1) by correction, zero-padding, FFT and interpolation."""
    ker = ESKernel.for_width(w)
    f = np.asarray(f, dtype=np.complex128)
    if f.ndim == 2:
        f = f[None]
    S, n, _ = f.shape
    nf = smooth_size(max(2 * n, 2 * w))
    x = np.ravel(x)
    y = np.ravel(y)
    ux = x * (nf / (2 * np.pi)) + nf / 2
    uy = y * (nf / (2 * np.pi)) + nf / 2
    lx, wx = _spread_weights(ux, ker)
    ly, wy = _spread_weights(uy, ker)
    k = modes(n)
    p = ker.correction(k, nf) * np.where(k % 2 == 0, 1.0, -1.0)
    out = np.zeros((S, x.size), dtype=np.complex128)
    for s in range(S):
        g = np.zeros((nf, nf), dtype=np.complex128)
        g[np.ix_(k % nf, k % nf)] = f[s] * p[:, None] * p[None, :]
        b = np.fft.fft2(g).ravel()
        acc = np.zeros(x.size, dtype=np.complex128)
        for a in range(w):
            iy = (ly + a) % nf
            for bb in range(w):
                ix = (lx + bb) % nf
                acc += b[iy * nf + ix] * (wy[:, a] * wx[:, bb])
        out[s] = acc
    return out


def _nufft1(c, x, y, nmodes, exact):
    return (nudft_type1 if exact else nufft_type1)(c, x, y, nmodes)


def _nufft2(f, x, y, exact):
    return (nudft_type2 if exact else nufft_type2)(f, x, y)


# ---------------------------------------------------------------------------
# A.3 / A.4 projectors, A.7 shifts, A.8 ramp
# ---------------------------------------------------------------------------


def centered_ray_dft(v: np.ndarray, sign: int) -> np.ndarray:
    """fftshift(FFT_sign(ifftshift(v))) along the last axis, unnormalised.

    ``src/gpu/fftshift.cuh:59-76`` (roll by -/+ floor(N/2)) and
    ``src/fft.h:101-151`` (cuFFT forward = exp(-2 pi i), inverse = exp(+2 pi i),
    both unnormalised); identical to numpy's shift/fft conventions.
    """
    n = v.shape[-1]
    u = np.fft.ifftshift(v, axes=-1)
    if sign < 0:
        u = np.fft.fft(u, axis=-1)
    else:
        u = np.fft.ifft(u, axis=-1) * n
    return np.fft.fftshift(u, axes=-1)


def ramlak(n: int) -> np.ndarray:
    """``src/gpu/ramlak.cu:19-29``: |z - floor(N/2)|/N (cutoff 1 never binds)."""
    f = np.abs(np.arange(n) - n // 2) / float(n)
    return np.where(f < 1.0, f, 0.0)


def project(img: np.ndarray, angles: np.ndarray, exact: bool = False):
    """``src/forward.cpp:33-52``: Re(fftshift(IFFT(ifftshift(A x))))/N^2.

    img [S, N, N] -> sinogram [S, P, N].
    """
    img = np.asarray(img, dtype=np.float64)
    if img.ndim == 2:
        img = img[None]
    S, n, _ = img.shape
    P = len(angles)
    x, y = polar_grid(angles, n)
    c = _nufft2(img, x, y, exact).reshape(S, P, n)
    out = centered_ray_dft(c, +1)
    return out.real / float(n * n)


def backproject(sino: np.ndarray, angles: np.ndarray, fbp: bool = False,
                exact: bool = False):
    """``src/backward.cpp:42-67``: Re(A^H(fftshift(FFT(ifftshift y)) R))/N^2.

    sinogram [S, P, N] -> image [S, N, N]; R = Ram-Lak ramp if ``fbp``.
    """
    sino = np.asarray(sino, dtype=np.float64)
    if sino.ndim == 2:
        sino = sino[None]
    S, P, n = sino.shape
    v = centered_ray_dft(sino.astype(np.complex128), -1)
    if fbp:
        v = v * ramlak(n)[None, None, :]
    x, y = polar_grid(angles, n)
    f = _nufft1(v.reshape(S, P * n), x, y, n, exact)
    return f.real / float(n * n)


# ---------------------------------------------------------------------------
# A.5, A.6 gradients and the Toeplitz operator
# ---------------------------------------------------------------------------


def gradient_nufft(x: np.ndarray, b: np.ndarray, angles: np.ndarray,
                   exact: bool = False):
    """``src/gradient.cpp:61-65``: Re(A^H A x)/N^3 - b."""
    x = np.asarray(x, dtype=np.float64)
    S, n, _ = x.shape
    gx, gy = polar_grid(angles, n)
    c = _nufft2(x, gx, gy, exact)
    f = _nufft1(c, gx, gy, n, exact)
    return f.real / float(n) ** 3 - b


def psf(angles: np.ndarray, n: int, exact: bool = False) -> np.ndarray:
    """``src/toeplitz.h:51-73``: PSF = Re(type-1(ones)) on (2n-1)^2 modes.

    Returns the real [2n-1, 2n-1] array with PSF(m) at index m + (n-1).
    """
    x, y = polar_grid(angles, n)
    ones = np.ones((1, x.size), dtype=np.complex128)
    return _nufft1(ones, x, y, 2 * n - 1, exact)[0].real


def toeplitz(x: np.ndarray, psf_arr: np.ndarray) -> np.ndarray:
    """``src/toeplitz.h:107-132`` (convolve1), restated.

    Zero-pad x (data top-left) to N1 x N1 (N1 = 2n-1), rfft2, multiply by
    rfft2(PSF), irfft2 (unnormalised), keep rows/cols n-1 .. 2n-2 and divide
    by n^3 * N1^2.  Equals sum_q x[q] PSF(a - q) / n^3.
    """
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    n = x.shape[-1]
    n1 = psf_arr.shape[-1]
    assert n1 == 2 * n - 1
    H = np.fft.rfft2(psf_arr)
    X = np.fft.rfft2(x, s=(n1, n1))
    y = np.fft.irfft2(X * H[None], s=(n1, n1))  # numpy divides by n1^2
    return y[:, n - 1:2 * n - 1, n - 1:2 * n - 1] / float(n) ** 3


def toeplitz_circulant(x: np.ndarray, psf_arr: np.ndarray, m: int):
    """Same operator on any FFT size m >= 2n-1 with the PSF placed circularly
    (PSF(m') at index m' mod m, zero elsewhere); the crop is the top-left n x n
    block.  Used to pin that the result does not depend on the FFT size."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    n = x.shape[-1]
    assert m >= 2 * n - 1
    circ = np.zeros((m, m))
    idx = np.arange(-(n - 1), n) % m
    circ[np.ix_(idx, idx)] = psf_arr
    H = np.fft.rfft2(circ)
    y = np.fft.irfft2(np.fft.rfft2(x, s=(m, m)) * H[None], s=(m, m))
    return y[:, :n, :n] / float(n) ** 3


def toeplitz_bruteforce(x: np.ndarray, psf_arr: np.ndarray) -> np.ndarray:
    """T x[a] = sum_q x[q] PSF(a-q) / n^3 by direct summation (tiny n)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    S, n, _ = x.shape
    out = np.zeros_like(x)
    for a0 in range(n):
        for a1 in range(n):
            blk = psf_arr[a0:a0 + n, a1:a1 + n][::-1, ::-1]
            out[:, a0, a1] = np.tensordot(x, blk, axes=([1, 2], [0, 1]))
    return out / float(n) ** 3


def gradient2(x: np.ndarray, b: np.ndarray, psf_arr: np.ndarray):
    """``src/gradient2.cpp:62``: g = T x - b."""
    return toeplitz(x, psf_arr) - b


def function_value(x: np.ndarray, sino: np.ndarray, angles: np.ndarray,
                   exact: bool = False) -> float:
    """``src/error.cpp:57-59``: sum (P x - y)^2."""
    d = project(x, angles, exact) - sino
    return float(np.sum(d * d))


def function_value2(x: np.ndarray, b: np.ndarray, psf_arr: np.ndarray,
                    sino_sq: float) -> float:
    """``src/error2.cpp:60-63,91``: ||y||^2 + sum(x . T x - 2 x . b)."""
    tx = toeplitz(x, psf_arr)
    return float(sino_sq + np.sum(x * tx) - 2.0 * np.sum(x * b))


# ---------------------------------------------------------------------------
# A.8 qGGMRF regularizer
# ---------------------------------------------------------------------------

MRF_Q = 2.0
MRF_C = 0.001
# src/gpu/potential_function.cuh:38-51 (26-neighbour weights, centre 0)
WEIGHTS = np.array([
    [[0.0302, 0.0370, 0.0302], [0.0370, 0.0523, 0.0370], [0.0302, 0.0370, 0.0302]],
    [[0.0370, 0.0523, 0.0370], [0.0523, 0.0000, 0.0523], [0.0370, 0.0523, 0.0370]],
    [[0.0302, 0.0370, 0.0302], [0.0370, 0.0523, 0.0370], [0.0302, 0.0370, 0.0302]],
])


def potential(delta, sigma: float, p: float):
    """``src/gpu/potential_function.cuh:64-68``: |d/s|^q / (c + |d/s|^(q-p))."""
    u = np.abs(delta) / sigma
    return u ** MRF_Q / (MRF_C + u ** (MRF_Q - p))


def d_potential(delta, sigma: float, p: float, dtype=np.float64):
    """``src/gpu/potential_function.cuh:90-108`` / ``test/tvd.cpp:24-36``.

    rho'(d) = sign(d) |d|^(q-1) / (sigma^q (c + t)) * (q - (q-p) t/(c+t)),
    t = |d|^(q-p) / sigma^(q-p); rho'(0) = 0.  ``dtype=np.float32`` evaluates
    it in single precision as the reference does even for T=double.
    """
    d = np.asarray(delta, dtype=dtype)
    sig = dtype(sigma)
    pp = dtype(p)
    q = dtype(MRF_Q)
    c = dtype(MRF_C)
    sq = sig ** q
    sqp = sig ** (q - pp)
    ad = np.abs(d)
    t1 = ad ** (q - pp) / sqp
    t2 = ad ** (q - dtype(1))
    t3 = c + t1
    mag = (t2 / (t3 * sq)) * (q - ((q - pp) * t1) / t3)
    return np.where(d > 0, mag, np.where(d < 0, -mag, dtype(0)))


def qggmrf_grad(x: np.ndarray, sigma: float, p: float,
                dtype=np.float64) -> np.ndarray:
    """``src/gpu/cuTV2.cu:62-73`` + clamp ``src/gpu/dev_memory.cuh:88-94``.

    Returns sum_d w_d rho'(x_v - x_clamp(v+d)) (the increment added to g).
    Replicate-at-edge clamp on slice, row and column axes.
    """
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    S, R, C = x.shape
    xp = np.pad(x, 1, mode="edge")
    acc = np.zeros_like(x)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                wgt = WEIGHTS[a, b, c]
                if wgt == 0.0:
                    continue
                nb = xp[a:a + S, b:b + R, c:c + C]
                acc += wgt * d_potential((x - nb).astype(dtype), sigma, p,
                                         dtype=dtype).astype(np.float64)
    return acc


def qggmrf_grad_loops(x: np.ndarray, sigma: float, p: float) -> np.ndarray:
    """``test/tvd.cpp:39-75`` (cpuTotalVar) restated loop-for-loop, in float32
    like the reference.  Pure Python: tiny volumes only."""
    x = np.asarray(x, dtype=np.float32)
    S, R, C = x.shape
    w32 = WEIGHTS.astype(np.float32)
    out = np.zeros_like(x)
    for i in range(S):
        for j in range(R):
            for k in range(C):
                u = x[i, j, k]
                v = np.float32(0.0)
                for z in range(3):
                    for yy in range(3):
                        for xx in range(3):
                            i1 = min(max(i + z - 1, 0), S - 1)
                            j1 = min(max(j + yy - 1, 0), R - 1)
                            k1 = min(max(k + xx - 1, 0), C - 1)
                            d = np.float32(u - x[i1, j1, k1])
                            v = np.float32(v + w32[z, yy, xx] *
                                           d_potential(d, sigma, p, np.float32))
                out[i, j, k] = v
    return out


def qggmrf_value(x: np.ndarray, sigma: float, p: float) -> float:
    """R(x) = 1/2 sum_v sum_d w_d rho(x_v - x_{v+d}) over in-bounds pairs.

    On interior voxels d R / d x_v equals :func:`qggmrf_grad` (the
    stencil is the exact gradient of this prior away from the clamp)."""
    x = np.asarray(x, dtype=np.float64)
    S, R, C = x.shape
    tot = 0.0
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                wgt = WEIGHTS[a + 1, b + 1, c + 1]
                if wgt == 0.0:
                    continue
                s0, s1 = max(0, -a), min(S, S - a)
                r0, r1 = max(0, -b), min(R, R - b)
                c0, c1 = max(0, -c), min(C, C - c)
                d = (x[s0:s1, r0:r1, c0:c1]
                     - x[s0 + a:s1 + a, r0 + b:r1 + b, c0 + c:c1 + c])
                tot += wgt * np.sum(potential(d, sigma, p))
    return 0.5 * tot


def qggmrf_grad_free(x: np.ndarray, sigma: float, p: float) -> np.ndarray:
    """Exact gradient of :func:`qggmrf_value` (free boundary): for each voxel,
    sum over the neighbours inside the volume of w_d rho'(x_v - x_{v+d}).
    In the interior it equals :func:`qggmrf_grad`; at the faces the pairs that
    leave the volume are dropped instead of clamped.  Used by the
    differentiable form (the reference has no such function)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    S, R, C = x.shape
    xp = np.pad(x, 1, mode="constant", constant_values=np.nan)
    acc = np.zeros_like(x)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                wgt = WEIGHTS[a, b, c]
                if wgt == 0.0:
                    continue
                nb = xp[a:a + S, b:b + R, c:c + C]
                inside = ~np.isnan(nb)
                d = np.where(inside, x - np.where(inside, nb, 0.0), 0.0)
                acc += np.where(inside, wgt * d_potential(d, sigma, p), 0.0)
    return acc


def tv_hessian_zero(sigma: float) -> float:
    """``src/gpu/tv_hessian.cu:36-48`` + ``potential_function.cuh:111-114``:
    sum_d w_d q / (sigma^q c), evaluated in float32 like the reference."""
    s = np.float32(sigma)
    d2 = np.float32(MRF_Q) / (s ** np.float32(MRF_Q) * np.float32(MRF_C))
    acc = np.float32(0.0)
    for wv in WEIGHTS.astype(np.float32).ravel():
        acc = np.float32(acc + wv * d2)
    return float(acc)


# ---------------------------------------------------------------------------
# A.11 pre / post processing
# ---------------------------------------------------------------------------


def pad_amount(nrays: int, center: float):
    """``src/preproc.cpp:80-85``: (npad, offset)."""
    # std::round: halves away from zero
    cen = int(math.copysign(math.floor(abs(center) + 0.5), center))
    off = (nrays + 1) // 2 - cen
    npad = int(0.42 * nrays) // 2
    if abs(off) > npad:
        npad = abs(off)
    return npad, off


def preproc(sino: np.ndarray, center: float) -> np.ndarray:
    """``src/preproc.cpp:62-65,73-108``: zero-pad each row by npad on both
    sides, then roll: out[z] = in[(z + off) mod N']."""
    sino = np.asarray(sino)
    S, P, N = sino.shape
    npad, off = pad_amount(N, center)
    n2 = N + 2 * npad
    padded = np.zeros((S, P, n2), dtype=sino.dtype)
    padded[:, :, npad:npad + N] = sino
    idx = (np.arange(n2) + off + n2) % n2
    return padded[:, :, idx]


def postproc(img: np.ndarray, npixels: int) -> np.ndarray:
    """``src/postproc.cpp:69-102``: crop [npad, npad+N) on both axes."""
    n2 = img.shape[-1]
    npad = (n2 - npixels) // 2
    return img[:, npad:n2 - npad, npad:n2 - npad]


def pad2d(img: np.ndarray, npad_total: int) -> np.ndarray:
    """``src/padding.cpp:37-91`` + ``src/gpu/padding.cu:119-150``: always
    symmetric (shift = npad_total/2) whatever the PadType argument."""
    S, R, C = img.shape
    shift = abs(npad_total) // 2
    out = np.zeros((S, R + abs(npad_total), C + abs(npad_total)), img.dtype)
    out[:, shift:shift + R, shift:shift + C] = img
    return out


# ---------------------------------------------------------------------------
# A.9 Optimizer::run2 and A.10 mbir2
# ---------------------------------------------------------------------------


@dataclass
class Run2Log:
    errors: List[float] = field(default_factory=list)
    xerrs: List[float] = field(default_factory=list)
    backtracks: int = 0
    steps: List[float] = field(default_factory=list)
    # optional per-attempt trace (run2 only): set to a list to record, for
    # every attempt, the scalars and arrays the control flow used
    attempts: Optional[List[dict]] = None


def run2(x0: np.ndarray, grad_fn: Callable, err_fn: Callable, num_iters: int,
         step: float, log: Optional[Run2Log] = None) -> np.ndarray:
    """``src/optimize.h:96-154`` restated line by line (tol/xtol unused)."""
    sol = np.array(x0, dtype=np.float64)
    x = sol.copy()
    t = 1.0
    tnew = 1.0
    step0 = step
    for _ in range(num_iters):
        while True:
            beta = tnew * (1.0 / t - 1.0)
            tnew = 0.5 * (math.sqrt(t ** 4 + 4.0 * t ** 2) - t ** 2)
            y = sol + (sol - x) * beta
            g = grad_fn(y)
            sol_prev = sol
            sol = y - g * step
            fx = err_fn(sol)
            fy = err_fn(y)
            gy = 0.5 * step * float(np.sum(g * g))
            if log is not None and log.attempts is not None:
                log.attempts.append(dict(
                    t=t, tnew=tnew, beta=beta, step=step, sol_prev=sol_prev,
                    x=x, y=y, g=g, cand=sol, accepted=not (fx > fy + gy)))
            if fx > fy + gy:
                step *= 0.9
                if log is not None:
                    log.backtracks += 1
            else:
                if log is not None:
                    log.steps.append(step)
                step = step0
                t = tnew
                xerr = float(np.sum((sol - x) ** 2))
                x = sol
                break
        e = err_fn(sol)
        if log is not None:
            log.errors.append(e)
            log.xerrs.append(math.sqrt(xerr))
    return sol


def run2_linear(x0: np.ndarray, T: Callable, b: np.ndarray, reg_grad: Callable,
                sino_norm: float, num_iters: int, step: float,
                log: Optional[Run2Log] = None) -> np.ndarray:
    """``run2`` with the mbir2 objective, rewritten to one application of the
    Toeplitz operator T per attempt (FP64).

    Same control flow as :func:`run2` (``src/optimize.h:96-154``): ``sol``
    takes the candidate on every attempt, ``x`` only on acceptance.  Uses
    T y = T sol + beta (T sol - T x), T cand = T y - s T g and
    err^2(v) = ||y||^2 + v.Tv - 2 v.b expanded in dot products.  Pinned to
    :func:`run2` by the tests (including retries).
    """
    sol = np.array(x0, dtype=np.float64)
    x = sol
    Tsol = T(sol)
    Tx = Tsol
    t = 1.0
    tnew = 1.0
    step0 = step
    for _ in range(num_iters):
        while True:
            beta = tnew * (1.0 / t - 1.0)
            tnew = 0.5 * (math.sqrt(t ** 4 + 4.0 * t ** 2) - t ** 2)
            y = sol + (sol - x) * beta
            Ty = Tsol + (Tsol - Tx) * beta
            g = Ty - b + reg_grad(y)
            Tg = T(g)
            s = step
            yTy = float(np.sum(y * Ty))
            yb = float(np.sum(y * b))
            fx = math.sqrt(sino_norm + yTy - 2 * s * float(np.sum(g * Ty))
                           + s * s * float(np.sum(g * Tg)) - 2 * yb
                           + 2 * s * float(np.sum(g * b)))
            fy = math.sqrt(sino_norm + yTy - 2 * yb)
            gy = 0.5 * s * float(np.sum(g * g))
            cand = y - s * g
            xerr = float(np.sum((cand - x) ** 2))
            sol = cand
            Tsol = Ty - s * Tg
            if fx > fy + gy:
                step *= 0.9
                if log is not None:
                    log.backtracks += 1
                continue
            if log is not None:
                log.steps.append(s)
            x = sol
            Tx = Tsol
            step = step0
            t = tnew
            break
        if log is not None:
            log.errors.append(fx)
            log.xerrs.append(math.sqrt(xerr))
    return sol


@dataclass
class MBIRSetup:
    sino2: np.ndarray     # normalised, padded sinogram [S, P, N']
    x0: np.ndarray        # padded FBP initial guess [S, N', N']
    sinoT: np.ndarray     # P^T sino2 [S, N', N']
    sino_norm: float      # sum(sino2^2)
    psf: np.ndarray       # [2N'-1, 2N'-1]
    lipschitz: float
    step: float
    nrays: int


def mbir2_setup(sino: np.ndarray, angles: np.ndarray, center: float,
                exact: bool = False) -> MBIRSetup:
    """``src/mbir2.cpp:46-113``: everything before ``Optimizer::run2``."""
    sino = np.asarray(sino, dtype=np.float64)
    if sino.ndim == 2:
        sino = sino[None]
    maxv = sino.max()
    if maxv == 0:
        raise ValueError("sinogram has ZERO max value")
    sino2 = sino / maxv
    x0 = backproject(sino2, angles, fbp=True, exact=exact)
    nrays = sino.shape[-1]
    sino2 = preproc(sino2, center)
    npad = sino2.shape[-1] - x0.shape[-1]
    x0 = pad2d(x0, npad)
    n2 = sino2.shape[-1]
    sinoT = backproject(sino2, angles, exact=exact)
    sino_norm = float(np.sum(sino2 * sino2))
    ps = psf(angles, n2, exact=exact)
    # L = max(gradient(ones, 0)); the qGGMRF Hessian is added on a device copy
    # only and never reaches the host array (src/gpu/tv_hessian.cu:53-65)
    ones = np.ones((1, n2, n2))
    L = float(np.max(toeplitz(ones, ps)))
    step = 1.0 / L
    if step > 1.0:
        step = 1.0
    return MBIRSetup(sino2, x0, sinoT, sino_norm, ps, L, step, nrays)


def power_lipschitz(psf_arr: np.ndarray, n: int, iters: int) -> float:
    """Largest eigenvalue of the Toeplitz operator by power iteration (FP64):
    v0 = ones / n, then ``iters`` times w = T v, lam = v.w / v.v,
    v = w / ||w||; returns the last lam.  (The mixed-precision form's option
    in place of the reference's max(T 1), src/mbir2.cpp:101-112.)"""
    v = np.full((1, n, n), 1.0 / n)
    lam = 0.0
    for _ in range(iters):
        w = toeplitz(v, psf_arr)
        lam = float(np.sum(v * w) / np.sum(v * v))
        v = w / math.sqrt(float(np.sum(w * w)))
    return lam


def mbir2(sino: np.ndarray, angles: np.ndarray, center: float, num_iters: int,
          sigma: float, p: float = 1.2, exact: bool = False,
          log: Optional[Run2Log] = None, setup: Optional[MBIRSetup] = None):
    """``src/mbir2.cpp:40-145``: normalise, FBP guess, pad, Toeplitz MBIR,
    crop.  Returns the [S, N, N] reconstruction (normalised units)."""
    st = setup if setup is not None else mbir2_setup(sino, angles, center,
                                                      exact)

    def grad_fn(x):
        return gradient2(x, st.sinoT, st.psf) + qggmrf_grad(x, sigma, p)

    def err_fn(x):
        return math.sqrt(function_value2(x, st.sinoT, st.psf, st.sino_norm))

    rec = run2(st.x0, grad_fn, err_fn, num_iters, st.step, log)
    return postproc(rec, st.nrays)


def mbir_nufft(x0: np.ndarray, sino: np.ndarray, angles: np.ndarray,
               center: float, num_iters: int, sigma: float, p: float = 1.2,
               exact: bool = False, log: Optional[Run2Log] = None):
    """``src/mbir.cpp:38-119``: the NUFFT normal-operator MBIR (no PSF).

    Normalise by (y - min) / (max - min) (x0 returned unchanged if
    max == min), pad the guess, b = P^T y, step = min(1, 1/max(Re(A^H A 1)
    / n^3)), then :func:`run2` with gradient Re(A^H A y)/n^3 - b + R'(y) and
    error ||P v - y||."""
    sino = np.asarray(sino, dtype=np.float64)
    if sino.ndim == 2:
        sino = sino[None]
    x0 = np.asarray(x0, dtype=np.float64)
    if x0.ndim == 2:
        x0 = x0[None]
    maxv, minv = sino.max(), sino.min()
    if maxv == minv:
        return x0
    sino2 = (sino - minv) / (maxv - minv)
    nrays = sino.shape[-1]
    sino2 = preproc(sino2, center)
    x0p = pad2d(x0, sino2.shape[-1] - x0.shape[-1])
    n2 = sino2.shape[-1]
    sinoT = backproject(sino2, angles, exact=exact)
    ones = np.ones((1, n2, n2))
    L = float(np.max(gradient_nufft(ones, np.zeros_like(ones), angles, exact)))
    step = min(1.0, 1.0 / L)

    def grad_fn(x):
        return gradient_nufft(x, sinoT, angles, exact) + qggmrf_grad(x, sigma, p)

    def err_fn(x):
        return math.sqrt(function_value(x, sino2, angles, exact))

    sol = run2(x0p, grad_fn, err_fn, num_iters, step, log)
    return postproc(sol, nrays)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel())
                 / max(np.linalg.norm(b.ravel()), 1e-300))
