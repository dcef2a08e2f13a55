"""The reference's Python entry points (tomocam/modeling.py:44-79,
tomocam/recon.py:5-40, tomocam/transform.py:5-56, README ``recon``).

``MBIR`` and ``recon`` keep the reference's defaults and checks: float32
casting, ``smoothness <= 0`` raises ``ValueError``, ``sigma = 1/smoothness``,
and the positional order ``(sino, angles, center, num_iters, sigma, tol,
xtol)`` into ``cTomocam.mbir``.  ``radon`` / ``radon_adj`` are the working
forms of tomocam/transform.py (whose calls into cTomocam do not match the
module's signatures): ``center`` / ``over_sample`` are accepted for
compatibility; ``radon`` has no centre parameter in cTomocam, ``radon_adj``
uses ``center`` as cTomocam.backproject does.
"""

from __future__ import annotations

import numpy as np

from . import cTomocam


def MBIR(sinogram, angles, center, num_iters=100, smoothness=0.01, tol=1.0E-04,
         xtol=1.0E-04, *, precision="fp32"):
    """Model-based iterative reconstruction (Toeplitz-gradient form).

    sinogram: [S, P, N] (or [P, N]) projection data; angles: [P] radians;
    center: rotation centre in detector pixels; smoothness > 0 (sigma =
    1/smoothness).  Returns the [S, N, N] reconstruction (in units of the
    sinogram normalised by its maximum, like the reference).

    ``precision="mixed"`` (keyword only; the reference is FP32 throughout)
    selects the mixed-precision form: FP16 gradient, FP32 spectrum,
    intermediates, FFTs and state, FP64 dots, T x refreshed in FP32 every
    25 iterations.
    """
    if sinogram.dtype != np.float32:
        sinogram = sinogram.astype(np.float32)
    if angles.dtype != np.float32:
        angles = angles.astype(np.float32)
    if smoothness <= 0:
        raise ValueError('smoothness value must be greater than 0')
    sigma = 1. / smoothness
    if precision != "fp32":
        return cTomocam.mbir(sinogram, angles, center, num_iters, sigma, tol,
                             xtol, precision=precision)
    return cTomocam.mbir(sinogram, angles, center, num_iters, sigma, tol, xtol)


def recon(sinogram, angles, center, num_iters=100, smoothness=0.01,
          tol=1.0E-04, xtol=1.0E-04, *, precision="fp32"):
    """README form ``recon(sino, theta, center)``; same as :func:`MBIR`."""
    return MBIR(sinogram, np.asarray(angles), center, num_iters, smoothness,
                tol, xtol, precision=precision)


def radon(volume, angles, center=None, over_sample=None):
    """Radon transform [S, N, N] -> [S, P, N] (cTomocam.radon)."""
    volume = np.asarray(volume)
    angles = np.asarray(angles)
    if volume.dtype != np.float32 and angles.dtype != np.float32:
        raise ValueError('input data-type must be single precision')
    return cTomocam.radon(volume, angles)


def radon_adj(sinogram, angles, center, over_sample=None):
    """Adjoint Radon transform with centring [S, P, N] -> [S, N, N]
    (cTomocam.backproject)."""
    sinogram = np.asarray(sinogram)
    angles = np.asarray(angles)
    if sinogram.dtype != np.float32 and angles.dtype != np.float32:
        raise ValueError('input data-type must be single precision')
    return cTomocam.backproject(sinogram, angles, center)
