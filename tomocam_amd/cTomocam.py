"""Drop-in for the reference's pybind11 module ``cTomocam``
(src/pyTomocam.cpp:137-152): ``mbir``, ``radon``, ``backproject`` and
``set_num_of_gpus`` with the same positional arguments and numpy in / out.

As with pybind11's ``array_t<float, c_style>`` arguments, inputs are converted
to C-contiguous float32; a 2-D input is one slice (src/pyTomocam.cpp:52-68)
and results are always 3-D ``[S, ...]`` (src/pyTomocam.cpp:70-79).  All work
runs in libtomocam_hip.so on the GPU; torch only stages host <-> device copies.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _stack3(a) -> np.ndarray:
    a = _f32(a)
    if a.ndim == 2:
        return a[None]
    if a.ndim != 3:
        raise ValueError("expected a 2-D or 3-D array")
    return a


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("tomocam_amd needs a visible GPU (HIP device)")
    _lib.lib()
    return torch


def set_num_of_gpus(num: int) -> None:
    """One GPU per process; ``num`` must be >= 1 (src/machine.h:48-54)."""
    _lib.call("tc_set_num_of_gpus", int(num))


def radon(imgstack, angs) -> np.ndarray:
    """Forward projection [S, N, N] -> [S, P, N] (src/pyTomocam.cpp:81-92)."""
    from . import gpu
    torch = _torch()
    img = _stack3(imgstack)
    out = gpu.project(torch.from_numpy(img).cuda(), _f32(angs))
    return out.cpu().numpy()


def backproject(sino, angs, cen) -> np.ndarray:
    """preproc + backproject + postproc (src/pyTomocam.cpp:94-114):
    [S, P, N] -> [S, N, N]."""
    from . import gpu
    torch = _torch()
    s = _stack3(sino)
    nrays = s.shape[-1]
    d = torch.from_numpy(s).cuda()
    padded = gpu.preproc(d, float(cen))
    img = gpu.backproject(padded, _f32(angs))
    return gpu.postproc(img, nrays).cpu().numpy()


def mbir(sino, angles, center, num_iters, sigma, tol, xtol, *,
         precision="fp32") -> np.ndarray:
    """tomocam::mbir2 (src/pyTomocam.cpp:116-134): [S, P, N] -> [S, N, N].
    ``precision="mixed"`` (keyword only, no counterpart in the reference)
    runs the mixed-precision form."""
    from . import gpu
    torch = _torch()
    s = _stack3(sino)
    ang = _f32(angles).ravel()
    if ang.size != s.shape[1]:
        raise ValueError(
            f"{ang.size} angles for a sinogram with {s.shape[1]} projections")
    out, _ = gpu.mbir(torch.from_numpy(s).cuda(), ang, float(center),
                      int(num_iters), float(sigma), float(tol), float(xtol),
                      verbose=True, precision=precision)
    return out.cpu().numpy()
