"""Toeplitz operator on the GPU vs the FP64 oracle (through the C-ABI).

Tolerance: relative L2 <= 2e-5 for T x (FP32 FFTs of length M ~ 2n), and
relative 1e-5 for the spectral x . T x.
"""

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


def _even_psf(n, seed):
    """Synthetic real PSF on (2n-1)^2 that is even (PSF(m) = PSF(-m)), as
    every polar-grid PSF is."""
    r = np.random.default_rng(seed).standard_normal((2 * n - 1, 2 * n - 1))
    return 0.5 * (r + r[::-1, ::-1])


@pytest.mark.parametrize("n,P,S", [(15, 7, 2), (16, 9, 3), (33, 20, 1),
                                    (48, 30, 2)])
def test_toeplitz_polar_psf(torch_cuda, n, P, S):
    torch = torch_cuda
    from tomocam_amd import gpu
    ang = np.arange(P) * np.pi / P
    ps = O.psf(ang, n, exact=True)
    x = np.random.default_rng(n).standard_normal((S, n, n))
    ref = O.toeplitz(x, ps)
    announce(f"toeplitz polar n={n} P={P} S={S}")
    psf = gpu.PointSpreadFunction(psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    xt = torch.tensor(x, dtype=torch.float32).cuda()
    out, xtx = psf.convolve(xt, with_dot=True)
    out = out.cpu().numpy()
    assert rel_l2(out, ref) < 2e-5
    # brute-force definition (tiny sizes)
    if n <= 16:
        assert rel_l2(out, O.toeplitz_bruteforce(x, ps)) < 2e-5
    # spectral dot x . T x
    assert abs(xtx - np.sum(x * ref)) <= 1e-5 * abs(np.sum(x * ref))


@pytest.mark.parametrize("n,S,M", [(181, 2, 378), (340, 2, 750),
                                   (362, 2, 768), (726, 1, 1458),
                                   (1024, 2, 2250), (1090, 2, 2304),
                                   (1454, 3, 2916), (2048, 1, 4500),
                                   (2180, 1, 4608), (2908, 1, 5832)])
def test_toeplitz_sizes(torch_cuda, n, S, M):
    """FFT length M of each compile-time plan (750, 768, 1458, 2250, 2304,
    2916, 4500, 4608, 5832) and of the generic kernels (378), pinned by
    the length rule's choice for the width."""
    torch = torch_cuda
    from tomocam_amd import gpu
    ps = _even_psf(n, 7)
    x = np.random.default_rng(n + 1).standard_normal((S, n, n))
    ref = O.toeplitz(x, ps)
    announce(f"toeplitz size n={n} S={S} M={M}")
    psf = gpu.PointSpreadFunction(psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    assert psf.fft_size == M
    out, xtx = psf.convolve(torch.tensor(x, dtype=torch.float32).cuda(),
                            with_dot=True)
    assert rel_l2(out.cpu().numpy(), ref) < 2e-5
    assert abs(xtx - np.sum(x * ref)) <= 1e-5 * np.sum(np.abs(x * ref))


def test_gradient2_function_value2(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    n, P, S = 32, 24, 2
    ang = np.arange(P) * np.pi / P
    ps = O.psf(ang, n, exact=True)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((S, n, n))
    b = rng.standard_normal((S, n, n))
    announce("gradient2/function_value2 n=32")
    psf = gpu.PointSpreadFunction(psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    xt = torch.tensor(x, dtype=torch.float32).cuda()
    bt = torch.tensor(b, dtype=torch.float32).cuda()
    g = psf.gradient2(xt, bt).cpu().numpy()
    assert rel_l2(g, O.gradient2(x, b, ps)) < 2e-5
    fv = psf.function_value2(xt, bt, 123.0)
    ref = O.function_value2(x, b, ps, 123.0)
    assert abs(fv - ref) <= 1e-5 * (abs(ref) + np.sum(np.abs(x * b)))


def test_psf_roundtrip_and_info(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    ps = _even_psf(20, 1)
    psf = gpu.PointSpreadFunction(psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    assert psf.n == 20
    assert psf.fft_size >= 39 and psf.fft_size % 2 == 0
    back = psf.array().cpu().numpy()
    assert np.array_equal(back, ps.astype(np.float32))


def test_toeplitz_group_autotuner(torch_cuda):
    """The slice-group autotuner picks one of its candidates, remembers it,
    and does not change T x (slices are independent; x . T x only changes
    in its FP64 summation order)."""
    torch = torch_cuda
    from tomocam_amd import gpu
    n, S = 362, 5
    ps = _even_psf(n, 11)
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    x = torch.tensor(np.random.default_rng(2).standard_normal((S, n, n)),
                     dtype=torch.float32).cuda()
    announce("toeplitz autotuner n=362 S=5")
    before, d0 = psf.convolve(x, with_dot=True)
    g = psf.tune(S)
    assert g in (2, 5)  # half and once the 144 MB rule (5 = all)
    assert psf.tune(S) == -1
    after, d1 = psf.convolve(x, with_dot=True)
    assert torch.equal(before, after)
    assert abs(d1 - d0) <= 1e-12 * abs(d0)
    assert psf.tune(1) == 0
    assert psf.tune(S, mixed=True) in (2, 5)
    small = gpu.PointSpreadFunction(
        psf_array=torch.tensor(_even_psf(181, 3), dtype=torch.float32).cuda())
    assert small.tune(S) == 0  # FFT length 378: no compile-time passes


def test_toeplitz_length_autotuner(torch_cuda):
    """The FFT-length autotuner at a width where two compile-time plans fit
    (n = 350: 750 = 10 15 5 and 768 = 24 8 4, both within 10 % of the
    smallest smooth length 700): it keeps one of them, later handles of the
    width get the same length, T x does not change beyond FP32 rounding
    (the product does not depend on the length), and a width with a single
    plan has nothing to tune."""
    torch = torch_cuda
    from tomocam_amd import gpu
    n, S = 350, 3
    ps = _even_psf(n, 21)
    x = np.random.default_rng(4).standard_normal((S, n, n))
    ref = O.toeplitz(x, ps)
    pst = torch.tensor(ps, dtype=torch.float32).cuda()
    xt = torch.tensor(x, dtype=torch.float32).cuda()
    psf = gpu.PointSpreadFunction(psf_array=pst)
    assert psf.fft_size == 768  # the static rule (768 measured cheaper)
    announce("toeplitz FFT-length autotuner n=350")
    M = psf.tune_length(S)
    assert M in (750, 768) and psf.fft_size == M
    assert rel_l2(psf.convolve(xt).cpu().numpy(), ref) < 2e-5
    again = gpu.PointSpreadFunction(psf_array=pst)
    assert again.fft_size == M
    assert again.tune_length(S) == M
    single = gpu.PointSpreadFunction(
        psf_array=torch.tensor(_even_psf(340, 5), dtype=torch.float32).cuda())
    assert single.tune_length(S) == 750  # only 750 fits n = 340
