"""The Python drop-in boundary: ``tomocam.MBIR`` / ``recon`` / ``radon`` /
``radon_adj`` (tomocam/modeling.py:44-79, tomocam/transform.py) and the
``cTomocam`` module (src/pyTomocam.cpp:137-152), host arrays in and out,
against the FP64 oracle; plus the ctypes stub of the host C entry point.

Tolerances: projectors 1e-5 relative L2 (exact NUDFT oracle), MBIR 1e-4.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu

N, P, S = 32, 24, 2


def _data(seed=31):
    ang = PH.angles(P)
    img = np.stack([PH.random_ellipses(N, seed + s) for s in range(S)])
    sino = O.project(img, ang.astype(np.float64), exact=True)
    sino = sino + np.random.default_rng(seed).normal(0, 0.01 * sino.max(),
                                                     sino.shape)
    return ang, img, sino


def test_cTomocam_mbir_matches_oracle(torch_cuda):
    from tomocam_amd import cTomocam
    ang, _, sino = _data()
    ref = O.mbir2(sino, ang.astype(np.float64), N / 2, 10, 10.0, exact=True)
    announce("cTomocam.mbir vs oracle")
    out = cTomocam.mbir(sino.astype(np.float32), ang, N / 2, 10, 10.0, 1e-4,
                        1e-4)
    assert out.shape == (S, N, N) and out.dtype == np.float32
    assert rel_l2(out, ref) < 1e-4
    # a 2-D sinogram is one slice (src/pyTomocam.cpp:52-68)
    one = cTomocam.mbir(sino[0].astype(np.float32), ang, N / 2, 10, 10.0,
                        1e-4, 1e-4)
    ref1 = O.mbir2(sino[:1], ang.astype(np.float64), N / 2, 10, 10.0,
                   exact=True)
    assert one.shape == (1, N, N)
    assert rel_l2(one, ref1) < 1e-4


def test_MBIR_and_recon_defaults(torch_cuda):
    """float64 inputs are cast; sigma = 1/smoothness (default 0.01 -> 100);
    recon is MBIR; smoothness <= 0 raises ValueError (modeling.py:70-73)."""
    import tomocam_amd as tomocam
    from tomocam_amd import cTomocam
    ang, _, sino = _data(37)
    announce("tomocam.MBIR / recon")
    a = tomocam.MBIR(sino, ang.astype(np.float64), N / 2, num_iters=8)
    b = cTomocam.mbir(sino.astype(np.float32), ang, N / 2, 8, 100.0, 1e-4,
                      1e-4)
    assert np.array_equal(a, b)
    c = tomocam.recon(sino, ang, N / 2, num_iters=8)
    assert np.array_equal(a, c)
    for bad in (0.0, -1.0):
        with pytest.raises(ValueError):
            tomocam.MBIR(sino, ang, N / 2, 2, smoothness=bad)


def test_radon_and_backproject(torch_cuda):
    import tomocam_amd as tomocam
    from tomocam_amd import cTomocam
    ang, img, _ = _data(41)
    a64 = ang.astype(np.float64)
    announce("cTomocam.radon / backproject")
    sino = cTomocam.radon(img.astype(np.float32), ang)
    assert rel_l2(sino, O.project(img, a64, exact=True)) < 1e-5
    cen = N / 2 + 1.5
    bp = cTomocam.backproject(sino, ang, cen)
    ref = O.postproc(O.backproject(O.preproc(sino.astype(np.float64), cen),
                                   a64, exact=True), N)
    assert bp.shape == (S, N, N)
    assert rel_l2(bp, ref) < 1e-5
    # the package's working radon / radon_adj forms
    assert np.array_equal(tomocam.radon(img.astype(np.float32), ang), sino)
    assert np.array_equal(tomocam.radon_adj(sino, ang, cen), bp)
    with pytest.raises(ValueError):
        tomocam.radon(img.astype(np.float64), ang.astype(np.float64))


def test_set_num_of_gpus():
    from tomocam_amd import cTomocam
    cTomocam.set_num_of_gpus(1)
    with pytest.raises(ValueError):
        cTomocam.set_num_of_gpus(0)


def test_host_c_entry_ctypes_stub(torch_cuda):
    """The minimal ctypes binding of tc_mbir2_host (host arrays, blocking)
    gives cTomocam.mbir's result."""
    from tomocam_amd import _lib, cTomocam
    lib = C.CDLL(_lib.lib_path())
    lib.tc_mbir2_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                  C.c_int, C.c_void_p, C.c_float, C.c_int,
                                  C.c_float, C.c_float, C.c_float, C.c_int]
    lib.tc_last_error.restype = C.c_char_p
    ang, _, sino = _data(43)
    s = np.ascontiguousarray(sino, np.float32)
    a = np.ascontiguousarray(ang, np.float32)
    out = np.empty((S, N, N), np.float32)
    announce("tc_mbir2_host through a plain ctypes stub")
    st = lib.tc_mbir2_host(s.ctypes.data, out.ctypes.data, S, P, N,
                           a.ctypes.data, N / 2, 8, 10.0, 1e-4, 1e-4, 0)
    assert st == 0, lib.tc_last_error()
    ref = cTomocam.mbir(s, a, N / 2, 8, 10.0, 1e-4, 1e-4)
    assert np.array_equal(out, ref)
