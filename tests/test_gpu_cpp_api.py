"""The tomocam:: C++ layer (include/tomocam/tomocam.hpp) end to end: the
driver tests/cpp/api_check (built by __graft_entry__.build()) runs on the GPU
with NPRandom(5489) inputs like the reference's test drivers; its outputs are
compared with the oracle here.  Tolerances as in the per-operator tests
(FP32 for T = float, 1e-10 for the T = double instantiations)."""

import os
import subprocess

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from conftest import ROOT, rel_l2, announce

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "api_check")


def test_cpp_interface(torch_cuda, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/api_check not built"
    announce("C++ tomocam:: interface")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    ld = lambda n, shape: np.fromfile(tmp_path / n, dtype=np.float32).reshape(shape)
    # the inputs are the NPRandom(5489) stream of test/timer.h
    rs = O.NPRandom()
    a = rs.rand(2 * 33 * 40).reshape(2, 33, 40).astype(np.float32)
    assert np.array_equal(ld("tv_in.f32", a.shape), a)
    tv = ld("tv_out.f32", a.shape)
    assert rel_l2(tv, O.qggmrf_grad(a, 10.0, 1.2)) < 1e-5
    N, P = 32, 24
    img = rs.rand(2 * N * N).reshape(2, N, N).astype(np.float32)
    assert np.array_equal(ld("img.f32", img.shape), img)
    ang = (np.arange(P) * np.pi / P).astype(np.float32).astype(np.float64)
    sino = ld("sino.f32", (2, P, N))
    assert rel_l2(sino, O.project(img, ang, exact=True)) < 1e-5
    bp = ld("bp.f32", (2, N, N))
    assert rel_l2(bp, O.backproject(sino, ang, exact=True)) < 1e-5
    ps = O.psf(ang, N, exact=True)
    x2 = rs.rand(2 * N * N).reshape(2, N, N).astype(np.float32)
    assert np.array_equal(ld("x2.f32", x2.shape), x2)
    g2 = ld("grad2.f32", (2, N, N))
    # T x2 and bp agree to ~5 %, so compare T x2 = g2 + bp (no
    # cancellation) and bound g2's error on the operands' scale
    Tx2 = O.toeplitz(x2, ps)
    assert rel_l2(g2.astype(np.float64) + bp, Tx2) < 1e-5
    assert np.abs(g2 - O.gradient2(x2, bp, ps)).max() <= 1e-5 * np.abs(Tx2).max()
    fv2 = float(ld("fv2.f32", (1,))[0])
    ref = O.function_value2(x2, bp, ps, float(np.sum(sino.astype(np.float64) ** 2)))
    assert abs(fv2 - ref) <= 1e-4 * abs(ref)
    rec = ld("mbir.f32", (2, N, N))
    ref = O.mbir2(sino, ang, N / 2, 10, 10.0, exact=True)
    assert rel_l2(rec, ref) < 1e-4
    assert "zero-sinogram threw: 1" in r.stdout

    # the T = double instantiations (FP64 kernels) on a fresh NPRandom
    # stream, against the oracle to FP64 tolerances
    assert "double forms ok" in r.stdout
    ld64 = lambda n, shape: np.fromfile(tmp_path / n, dtype=np.float64).reshape(shape)
    rs = O.NPRandom()
    a = rs.rand(2 * 33 * 40).reshape(2, 33, 40)
    assert rel_l2(ld64("tv_out.f64", a.shape), O.qggmrf_grad(a, 10.0, 1.2)) < 1e-12
    img = rs.rand(2 * N * N).reshape(2, N, N)
    ang = np.arange(P) * np.pi / P
    sino = ld64("sino.f64", (2, P, N))
    assert rel_l2(sino, O.project(img, ang, exact=True)) < 1e-10
    bp = ld64("bp.f64", (2, N, N))
    assert rel_l2(bp, O.backproject(sino, ang, exact=True)) < 1e-10
    ps = O.psf(ang, N, exact=True)
    x2 = rs.rand(2 * N * N).reshape(2, N, N)
    g2 = ld64("grad2.f64", (2, N, N))
    assert rel_l2(g2 + bp, O.toeplitz(x2, ps)) < 1e-10
    fv2 = float(ld64("fv2.f64", (1,))[0])
    ref = O.function_value2(x2, bp, ps, float(np.sum(sino ** 2)))
    assert abs(fv2 - ref) <= 1e-10 * abs(ref)
    xe = float(ld64("xerror.f64", (1,))[0])
    assert abs(xe - np.sum((img - x2) ** 2)) <= 1e-13 * xe
    pre = O.preproc(sino, 13.0)
    assert np.array_equal(ld64("preproc.f64", pre.shape), pre)
    rec = ld64("mbir.f64", (2, N, N))
    assert rel_l2(rec, O.mbir2(sino, ang, N / 2, 10, 10.0, exact=True)) < 1e-10
    rec1 = ld64("mbir_nufft.f64", (2, N, N))
    ref1 = O.mbir_nufft(np.zeros((2, N, N)), sino, ang, N / 2, 5, 10.0,
                        exact=True)
    assert rel_l2(rec1, ref1) < 1e-10
