"""Synthetic phantoms and geometry for the MBIR workloads (numpy only).

* ``shepp_logan``: the 10-ellipse modified Shepp-Logan phantom (Toft's
  intensities, values in [0, 1]).
* ``random_ellipses``: seeded sums of random ellipses (centres U(-0.5, 0.5)^2,
  semi-axes U[0.02, 0.4], rotation U[0, pi), density U[0, 1], overlaps add).
* ``angles``: theta_i = i pi / P.

Pixel centres are at (k - N/2) / (N/2) on both axes, so the image spans
[-1, 1).
"""

from __future__ import annotations

import numpy as np

# (intensity, semi-axis a, semi-axis b, x0, y0, rotation in degrees)
_SHEPP_LOGAN_TOFT = (
    (1.0, 0.69, 0.92, 0.0, 0.0, 0.0),
    (-0.8, 0.6624, 0.8740, 0.0, -0.0184, 0.0),
    (-0.2, 0.1100, 0.3100, 0.22, 0.0, -18.0),
    (-0.2, 0.1600, 0.4100, -0.22, 0.0, 18.0),
    (0.1, 0.2100, 0.2500, 0.0, 0.35, 0.0),
    (0.1, 0.0460, 0.0460, 0.0, 0.1, 0.0),
    (0.1, 0.0460, 0.0460, 0.0, -0.1, 0.0),
    (0.1, 0.0460, 0.0230, -0.08, -0.605, 0.0),
    (0.1, 0.0230, 0.0230, 0.0, -0.606, 0.0),
    (0.1, 0.0230, 0.0460, 0.06, -0.605, 0.0),
)


def _pixel_coords(n: int):
    t = (np.arange(n, dtype=np.float64) - n / 2.0) / (n / 2.0)
    yy, xx = np.meshgrid(t, t, indexing="ij")
    return xx, yy


def _raster(n: int, ellipses) -> np.ndarray:
    xx, yy = _pixel_coords(n)
    img = np.zeros((n, n), dtype=np.float64)
    for rho, a, b, x0, y0, phi in ellipses:
        c, s = np.cos(phi), np.sin(phi)
        dx = xx - x0
        dy = yy - y0
        u = dx * c + dy * s
        v = -dx * s + dy * c
        img[(u / a) ** 2 + (v / b) ** 2 <= 1.0] += rho
    return img


def shepp_logan(n: int) -> np.ndarray:
    """Modified (Toft) Shepp-Logan phantom, [n, n] float64 in [0, 1]."""
    ell = [(r, a, b, x0, y0, np.deg2rad(ph))
           for r, a, b, x0, y0, ph in _SHEPP_LOGAN_TOFT]
    return np.clip(_raster(n, ell), 0.0, 1.0)


def random_ellipses(n: int, seed: int, count: int = 30) -> np.ndarray:
    """[n, n] float64 sum of ``count`` random ellipses from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    ell = []
    for _ in range(count):
        x0, y0 = rng.uniform(-0.5, 0.5, size=2)
        a, b = rng.uniform(0.02, 0.4, size=2)
        phi = rng.uniform(0.0, np.pi)
        rho = rng.uniform(0.0, 1.0)
        ell.append((rho, a, b, x0, y0, phi))
    return _raster(n, ell)


def phantom_stack(n: int, nslices: int, kind: str = "ellipses",
                  seed_group: int = 1) -> np.ndarray:
    """[nslices, n, n] float32 stack; ellipse slices use seed
    1000*seed_group + slice."""
    if kind == "shepp":
        sl = shepp_logan(n)
        return np.repeat(sl[None], nslices, axis=0).astype(np.float32)
    out = np.empty((nslices, n, n), dtype=np.float32)
    for s in range(nslices):
        out[s] = random_ellipses(n, 1000 * seed_group + s)
    return out


def angles(nproj: int) -> np.ndarray:
    """theta_i = i pi / P, float32."""
    return (np.arange(nproj, dtype=np.float64) * np.pi / nproj).astype(
        np.float32)


def add_noise(sino: np.ndarray, rel_sd: float, seed: int) -> np.ndarray:
    """Additive Gaussian noise with sd = rel_sd * max(sino)."""
    rng = np.random.default_rng(seed)
    sd = rel_sd * float(np.max(sino))
    return (sino + rng.normal(0.0, sd, size=sino.shape)).astype(np.float32)
