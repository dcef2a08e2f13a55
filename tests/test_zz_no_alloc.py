"""tc_toeplitz_apply_noalloc (the allocation-free Toeplitz product) against
tc_toeplitz_apply, through the C-ABI, on two calls in a row (the second
with fewer slices, then a smaller width), and its size limit.

Tolerance: T x equal bit for bit (slices are independent and the kernels
are the same; only the slice-group size differs); x . T x (FP64 partial
sums, whose grouping follows the slice groups) within 1e-12 relative.
"""

import ctypes as C

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _even_psf(n, seed):
    r = np.random.default_rng(seed).standard_normal((2 * n - 1, 2 * n - 1))
    return 0.5 * (r + r[::-1, ::-1])


def _both(torch, psf, x):
    """(out, xTx) of the allocating and of the allocation-free entry point
    on the current stream."""
    from tomocam_amd import _lib
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    S = x.shape[0]
    res = []
    for name in ("tc_toeplitz_apply", "tc_toeplitz_apply_noalloc"):
        out = torch.full_like(x, float("nan"))
        d = C.c_double()
        _lib.call(name, psf.handle, C.c_void_p(x.data_ptr()),
                  C.c_void_p(out.data_ptr()), S, C.byref(d), s)
        res.append((out, d.value))
    return res


@pytest.mark.parametrize("n,slices", [(1454, (6, 2)), (181, (5, 3))])
def test_noalloc_equals_allocating_apply(torch_cuda, n, slices):
    """Width 1454 (compile-time FFT plan 2916, four slices per group fit
    the scratch) and 181 (runtime-plan FFT 378): two calls in a row, the
    second on fewer slices."""
    torch = torch_cuda
    from tomocam_amd import gpu
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(_even_psf(n, 5), dtype=torch.float32).cuda())
    rng = np.random.default_rng(n)
    for S in slices:
        x = torch.tensor(rng.standard_normal((S, n, n)),
                         dtype=torch.float32).cuda()
        announce(f"toeplitz noalloc n={n} S={S}")
        (a, da), (b, db) = _both(torch, psf, x)
        assert torch.equal(a, b), f"n={n} S={S}"
        assert abs(da - db) <= 1e-12 * abs(da)


def test_noalloc_smaller_width_after_larger(torch_cuda):
    """The scratch keeps the previous call's contents: a width-362 product
    right after a width-2908 one still equals the allocating product."""
    torch = torch_cuda
    from tomocam_amd import gpu
    for n, S in ((2908, 1), (362, 3)):
        psf = gpu.PointSpreadFunction(
            psf_array=torch.tensor(_even_psf(n, 9), dtype=torch.float32).cuda())
        x = torch.tensor(np.random.default_rng(n).standard_normal((S, n, n)),
                         dtype=torch.float32).cuda()
        announce(f"toeplitz noalloc n={n} S={S}")
        (a, da), (b, db) = _both(torch, psf, x)
        assert torch.equal(a, b), f"n={n}"
        assert abs(da - db) <= 1e-12 * abs(da)


def test_noalloc_rejects_oversized_request(torch_cuda):
    """More dot partials than the scratch holds: a clear error before any
    work (the output stays untouched)."""
    torch = torch_cuda
    from tomocam_amd import gpu, _lib
    n = 1454
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(_even_psf(n, 3), dtype=torch.float32).cuda())
    x = torch.zeros((1, n, n), dtype=torch.float32, device="cuda")
    out = torch.zeros_like(x)
    d = C.c_double()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    announce("toeplitz noalloc oversized request")
    # 10000 slices x 1459 partials x 8 bytes > 72 MiB: rejected up front
    with pytest.raises(ValueError, match="no-alloc Toeplitz apply: needs"):
        _lib.call("tc_toeplitz_apply_noalloc", psf.handle,
                  C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), 10000,
                  C.byref(d), s)
    torch.cuda.synchronize()
    assert bool((out == 0).all())
