"""Device-resident Optimizer::run2 vs the FP64 oracle restatement.

The set-up (x0, sinoT, PSF, ||y||^2, step) comes from the oracle so this
pins the iteration alone.  Tolerance: relative L2 <= 1e-4 on the solution
(the parity bound) and 1e-5 on the per-iteration error history.
"""

import math

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


def make_setup(N=32, P=24, S=3, seed=11):
    ang = PH.angles(P).astype(np.float64)
    img = np.stack([PH.random_ellipses(N, seed + s) for s in range(S)])
    sino = O.project(img, ang, exact=True)
    sino = sino + np.random.default_rng(seed).normal(0, 0.01 * sino.max(),
                                                     sino.shape)
    return O.mbir2_setup(sino, ang, N / 2, exact=True)


@pytest.fixture(scope="module")
def small_setup():
    return make_setup()


def oracle_run2(st, x0, iters, step, sigma):
    log = O.Run2Log()

    def grad_fn(x):
        return O.gradient2(x, st.sinoT, st.psf) + O.qggmrf_grad(x, sigma, 1.2)

    def err_fn(x):
        return math.sqrt(O.function_value2(x, st.sinoT, st.psf, st.sino_norm))

    return O.run2(x0, grad_fn, err_fn, iters, step, log), log


def gpu_run2(torch, st, x0, iters, step, sigma, refresh=10):
    from tomocam_amd import gpu
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(st.psf, dtype=torch.float32).cuda())
    x0t = torch.tensor(x0, dtype=torch.float32).cuda()
    b = torch.tensor(st.sinoT, dtype=torch.float32).cuda()
    return gpu.run2(psf, x0t, b, iters, step, sigma, 1.2, st.sino_norm,
                    refresh_every=refresh)


@pytest.mark.parametrize("sigma", [10.0, 100.0])
def test_run2_matches_oracle(torch_cuda, small_setup, sigma):
    st = small_setup
    iters = 30
    ref, log = oracle_run2(st, st.x0, iters, st.step, sigma)
    announce(f"run2 sigma={sigma}")
    out, err, xerr, stats = gpu_run2(torch_cuda, st, st.x0, iters, st.step,
                                     sigma)
    assert stats["iterations"] == iters
    assert stats["backtracks"] == log.backtracks
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    np.testing.assert_allclose(err, log.errors, rtol=1e-5)
    np.testing.assert_allclose(xerr, log.xerrs, rtol=1e-3, atol=1e-6)


def test_run2_backtracking(torch_cuda, small_setup):
    """A smooth perturbation of a converged point with step 6 makes the
    reference reject 4 candidates in two iterations; a rejected candidate
    still becomes `sol` (src/optimize.h:120,127-128)."""
    st = small_setup
    sigma = 1e4
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, sigma, 1.2)
    xs = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, 150, 3.0)
    x0 = xs + 1e-3
    ref, log = oracle_run2(st, x0, 2, 6.0, sigma)
    assert log.backtracks == 4
    announce("run2 backtracking")
    out, err, _, stats = gpu_run2(torch_cuda, st, x0, 2, 6.0, sigma)
    assert stats["backtracks"] == log.backtracks
    assert stats["attempts"] == 2 + log.backtracks
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    np.testing.assert_allclose(err, log.errors, rtol=1e-4)


@pytest.mark.parametrize("refresh", [0, 1, 10])
def test_run2_refresh(torch_cuda, small_setup, refresh):
    """With or without recomputing T(sol) the result meets the bound."""
    st = small_setup
    ref, _ = oracle_run2(st, st.x0, 40, st.step, 10.0)
    announce(f"run2 refresh={refresh}")
    out = gpu_run2(torch_cuda, st, st.x0, 40, st.step, 10.0, refresh)[0]
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4


def test_run2_divergent_sigma_first_iterations(torch_cuda, small_setup):
    """sigma = 0.1 (the benchmark's regulariser) makes the reference
    algorithm diverge; the early trajectory still matches."""
    st = small_setup
    ref, log = oracle_run2(st, st.x0, 5, st.step, 0.1)
    announce("run2 sigma=0.1")
    out, err, _, _ = gpu_run2(torch_cuda, st, st.x0, 5, st.step, 0.1)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    np.testing.assert_allclose(err, log.errors, rtol=1e-4)




def gaussian_psf(n):
    """A Gaussian-sum PSF (positive spectrum: T is positive definite),
    scaled so that max(T 1) is about 1."""
    m = np.arange(-(n - 1), n, dtype=np.float64)
    r2 = m[:, None] ** 2 + m[None, :] ** 2
    ps = np.exp(-r2 / (2 * 2.5 ** 2)) + 0.02 * np.exp(-r2 / (2 * 60.0 ** 2))
    return ps * (n ** 3 / ps.sum())


# (n, slices, step * L, retries).  At these widths the FFT length has
# compile-time plans and the step is committed inside the Toeplitz passes.
# The optimum is 0 (b = 0, ||y||^2 = 1e-4), so no dot product cancels, and
# the ratios (fx - fy) / gy of all attempts stay >= 36 % away from the
# accept threshold 1: FP32 cannot flip a decision.
@pytest.mark.parametrize("n,S,mult,retries", [(362, 2, 7.0, 2),
                                              (1454, 1, 4.5, 1)])
def test_run2_fused_commit_with_retries(torch_cuda, n, S, mult, retries):
    """x0 = 1e-3 cos x cos y with a long step: the reference rejects the
    first candidates (retries with x != sol), then accepts.  The fused
    passes must also match the separate commit kernel bit for bit."""
    import os
    from tomocam_amd import gpu
    sigma, iters, sino_norm = 1e4, 3, 1e-4
    ps = gaussian_psf(n)
    L = float(np.max(O.toeplitz(np.ones((1, n, n)), ps)))
    u = (np.arange(n) + 0.5) / n
    x0 = 1e-3 * (np.cos(np.pi * u)[None, :, None] *
                 np.cos(np.pi * u)[None, None, :]) * np.ones((S, 1, 1))
    b = np.zeros_like(x0)
    step = mult / L
    log = O.Run2Log()
    ref = O.run2_linear(x0, lambda v: O.toeplitz(v, ps), b,
                        lambda v: O.qggmrf_grad(v, sigma, 1.2), sino_norm,
                        iters, step, log)
    assert log.backtracks == retries
    announce(f"run2 fused commit n={n} S={S} ({retries} retries)")
    t = torch_cuda
    psf = gpu.PointSpreadFunction(
        psf_array=t.tensor(ps, dtype=t.float32).cuda())
    x0t = t.tensor(x0, dtype=t.float32).cuda()
    bt = t.tensor(b, dtype=t.float32).cuda()
    out, err, _, stats = gpu.run2(psf, x0t, bt, iters, step, sigma, 1.2,
                                  sino_norm)
    assert stats["backtracks"] == retries
    assert stats["attempts"] == iters + retries
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    np.testing.assert_allclose(err, log.errors, rtol=1e-4)
    os.environ["TOMOCAM_HIP_FUSED_COMMIT"] = "0"
    try:
        out2, err2, _, stats2 = gpu.run2(psf, x0t, bt, iters, step, sigma,
                                         1.2, sino_norm)
    finally:
        del os.environ["TOMOCAM_HIP_FUSED_COMMIT"]
    assert stats2["backtracks"] == retries
    assert t.equal(out, out2)
    np.testing.assert_array_equal(err, err2)


def test_run2_slice_groups_and_two_streams(torch_cuda, monkeypatch):
    """The compile-time passes in several slice groups (5 slices in groups of
    2: 2 + 2 + 1, or of 1), on one stream and round-robin over two or three,
    against one group and against the FP64 oracle (padded width 362, FFT
    length 768)."""
    torch = torch_cuda
    ang = PH.angles(60).astype(np.float64)
    img = np.stack([PH.random_ellipses(256, 500 + s) for s in range(5)])
    sino = O.project(img, ang)
    st = O.mbir2_setup(sino, ang, 128.0)
    iters, sigma = 15, 100.0
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, sigma, 1.2)
    ref = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, iters, st.step)
    monkeypatch.setenv("TOMOCAM_HIP_AUTOTUNE", "0")
    outs = {}
    for groups, streams in (("0", "1"), ("2", "1"), ("2", "2"), ("1", "3"),
                            ("2", "3")):
        monkeypatch.setenv("TOMOCAM_HIP_TOEPLITZ_SLICES", groups)
        monkeypatch.setenv("TOMOCAM_HIP_TOEPLITZ_STREAMS", streams)
        announce(f"run2 slice groups {groups} streams {streams}")
        out = gpu_run2(torch, st, st.x0, iters, st.step, sigma)[0]
        outs[(groups, streams)] = out.cpu().numpy()
        assert rel_l2(outs[(groups, streams)], ref) < 1e-4
    one = outs[("0", "1")]
    assert rel_l2(outs[("2", "1")], one) < 1e-6
    assert rel_l2(outs[("2", "2")], one) < 1e-6
    assert rel_l2(outs[("1", "3")], one) < 1e-6
    assert rel_l2(outs[("2", "3")], one) < 1e-6
