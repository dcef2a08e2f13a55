"""Compile-time line FFT plans (fft_lines_ct.hip) against the runtime-plan
line kernel, which the exact-NUDFT projector tests pin.

For every planned length L up to 8748, a projection and a backprojection
of width N = L / 2 use an L x L fine grid (rows and pruned columns, both
transform directions; the projectors' ray DFT limits N to 4374); the PSF of
width n uses a fine grid of the smallest planned length >= 2 (2n - 1)
(11664 at n = 2908) and a Toeplitz spectrum FFT.  Each result with the
planned kernels must match the same call on the runtime-plan kernel
(TOMOCAM_HIP_CT_LINES=0) to relL2 <= 2e-6 (two FP32 FFT orderings).  The
256^2 projector test in test_gpu_nufft.py compares the planned path with
the FP64 oracle directly."""

import os

import numpy as np
import pytest

from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu

LENGTHS = [512, 768, 1024, 1458, 1536, 2048, 2304, 2916, 3072, 4374, 4608,
           5832, 6144, 8748]


def _both(fn):
    """fn() with the planned kernels, then with the runtime-plan kernel."""
    os.environ.pop("TOMOCAM_HIP_CT_LINES", None)
    a = fn()
    os.environ["TOMOCAM_HIP_CT_LINES"] = "0"
    try:
        b = fn()
    finally:
        os.environ.pop("TOMOCAM_HIP_CT_LINES", None)
    return a, b


@pytest.mark.parametrize("L", LENGTHS)
def test_projectors_on_planned_fine_grid(torch_cuda, L):
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P = L // 2, 8
    announce(f"planned line FFT L={L} (projectors at N={N})")
    ang = PH.angles(P)
    g = torch.Generator(device="cuda")
    g.manual_seed(L)
    x = torch.rand((1, N, N), device="cuda", generator=g)
    s = torch.rand((1, P, N), device="cuda", generator=g)
    y_ct, y_rt = _both(lambda: gpu.project(x, ang).double().cpu().numpy())
    assert rel_l2(y_ct, y_rt) <= 2e-6
    b_ct, b_rt = _both(lambda: gpu.backproject(s, ang).double().cpu().numpy())
    assert rel_l2(b_ct, b_rt) <= 2e-6


@pytest.mark.parametrize("n", [362, 1454, 2908])
def test_psf_on_planned_grids(torch_cuda, n):
    torch = torch_cuda
    from tomocam_amd import gpu
    announce(f"planned line FFT: PSF n={n}")
    ang = PH.angles(16)

    def run():
        psf = gpu.PointSpreadFunction(ang, n)
        x = torch.ones((1, n, n), device="cuda")
        return (psf.array().double().cpu().numpy(),
                psf.convolve(x).double().cpu().numpy())

    (p_ct, t_ct), (p_rt, t_rt) = _both(run)
    assert rel_l2(p_ct, p_rt) <= 2e-6
    assert rel_l2(t_ct, t_rt) <= 2e-6
