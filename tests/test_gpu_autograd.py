"""Differentiable form (torch.autograd) vs FP64 finite differences and the
FP64 oracle's P^T (P x - y).

- The backward of the projector must be its exact adjoint: the autograd
  gradient of 0.5 ||P x - y||^2 equals P^T (P x - y) (relL2 <= 1e-5 against
  the exact-NUDFT oracle).
- The free-boundary qGGMRF value and gradient match the oracle (value to
  1e-6 relative, gradient relL2 <= 1e-5), whose gradient is itself pinned
  to central finite differences of its value (tests/test_oracle.py).
- Directional derivatives of the full loss 0.5 ||P x - y||^2 + beta R(x)
  along 64 random probe directions match FP64 central finite differences
  (eps = 1e-6) of the oracle loss: each within 1e-4 |grad| |d| and relL2
  over the probes <= 1e-4.
- The differentiation workload's shape (8 x 512^2, 360 angles, beta = 1e-3): the gradient
  matches the oracle on the first two slices (relL2 <= 1e-4; the prior
  term uses the whole volume), and 64 random directional derivatives match
  FP64 central finite differences (eps = 1e-6) of the loss evaluated by the
  FP64 kernels, which are themselves pinned to the oracle to 1e-10
  (tests/test_gpu_fp64.py): within 1e-4 |grad| for the FP32 gradient and
  1e-6 |grad| for the FP64 one.
"""

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


def _loss_grad(torch, x, y, ang, beta, sigma, dtype):
    from tomocam_amd import autograd as AG
    xt = torch.tensor(x, dtype=dtype, device="cuda", requires_grad=True)
    yt = torch.tensor(y, dtype=dtype, device="cuda")
    loss = AG.mbir_loss(xt, yt, ang, beta=beta, sigma=sigma, p=1.2)
    loss.backward()
    return (float(loss.detach()),
            xt.grad.detach().cpu().numpy().astype(np.float64))


def _oracle_loss(x, y, ang, beta, sigma):
    r = O.project(x, ang, exact=True) - y
    return 0.5 * float(np.sum(r * r)) + beta * O.qggmrf_value(x, sigma, 1.2)


@pytest.mark.parametrize("shape", [(2, 48, 40), (1, 33, 20)])
def test_projector_backward_is_adjoint(torch_cuda, shape):
    S, N, P = shape
    ang = PH.angles(P).astype(np.float64)
    x = np.stack([PH.random_ellipses(N, 60 + s) for s in range(S)])
    y = O.project(np.stack([PH.random_ellipses(N, 80 + s) for s in range(S)]),
                  ang, exact=True)
    announce(f"autograd data term {shape}")
    _, g = _loss_grad(torch_cuda, x, y, ang, 0.0, 0.1, torch_cuda.float32)
    ref = O.backproject(O.project(x, ang, exact=True) - y, ang, exact=True)
    assert rel_l2(g, ref) < 1e-5


@pytest.mark.parametrize("shape", [(3, 17, 65), (2, 40, 40), (1, 5, 7)])
@pytest.mark.parametrize("sigma", [0.1, 10.0])
def test_qggmrf_free_matches_oracle(torch_cuda, shape, sigma):
    from tomocam_amd import gpu
    x = np.random.default_rng(7).random(shape)
    announce(f"free qGGMRF {shape} sigma={sigma}")
    xt = torch_cuda.tensor(x, dtype=torch_cuda.float32, device="cuda")
    g = torch_cuda.zeros_like(xt)
    gpu.qggmrf_grad_free(xt, g, sigma, 1.2)
    val = gpu.qggmrf_value_free(xt, sigma, 1.2)
    x32 = x.astype(np.float32).astype(np.float64)
    assert rel_l2(g.cpu().numpy(), O.qggmrf_grad_free(x32, sigma, 1.2)) < 1e-5
    ref = O.qggmrf_value(x32, sigma, 1.2)
    assert abs(val - ref) <= 1e-6 * abs(ref)


def test_loss_directional_derivatives_vs_fd(torch_cuda):
    S, N, P, beta, sigma = 2, 40, 36, 1e-3, 0.1
    ang = PH.angles(P).astype(np.float64)
    x = np.stack([PH.random_ellipses(N, 90 + s) for s in range(S)])
    x = x + 0.05 * np.random.default_rng(1).random(x.shape)
    y = O.project(np.stack([PH.random_ellipses(N, 95 + s) for s in range(S)]),
                  ang, exact=True)
    x = x.astype(np.float32).astype(np.float64)
    announce("autograd loss vs 64 FP64 finite differences")
    _, g = _loss_grad(torch_cuda, x, y, ang, beta, sigma, torch_cuda.float32)
    rng = np.random.default_rng(2)
    eps = 1e-6
    got, fd = [], []
    for _ in range(64):
        d = rng.standard_normal(x.shape)
        d /= np.linalg.norm(d)
        lp = _oracle_loss(x + eps * d, y, ang, beta, sigma)
        lm = _oracle_loss(x - eps * d, y, ang, beta, sigma)
        fd.append((lp - lm) / (2 * eps))
        got.append(float(np.sum(g * d)))
    got, fd = np.array(got), np.array(fd)
    scale = np.linalg.norm(g)
    assert np.all(np.abs(got - fd) <= 1e-4 * scale)
    assert rel_l2(got, fd) < 1e-4


def test_float64_input(torch_cuda):
    """float64 tensors take the FP64 kernels (to 1e-10 of the oracle),
    float32 ones the FP32 kernels (to FP32 rounding of the same values)."""
    S, N, P = 1, 32, 24
    ang = PH.angles(P).astype(np.float64)
    x = PH.random_ellipses(N, 3)[None]
    y = O.project(PH.random_ellipses(N, 4)[None], ang, exact=True)
    announce("autograd float64 tensors")
    l32, g32 = _loss_grad(torch_cuda, x, y, ang, 1e-3, 0.1, torch_cuda.float32)
    l64, g64 = _loss_grad(torch_cuda, x, y, ang, 1e-3, 0.1, torch_cuda.float64)
    r = O.project(x, ang, exact=True) - y
    ref = O.backproject(r, ang, exact=True) + \
        1e-3 * O.qggmrf_grad_free(x, 0.1, 1.2)
    assert rel_l2(g64, ref) < 1e-10
    assert abs(l64 - _oracle_loss(x, y, ang, 1e-3, 0.1)) <= 1e-10 * abs(l64)
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    assert rel_l2(g32, g64) < 1e-5


def test_config2_shape(torch_cuda):
    """8 slices of 512^2, 360 angles, beta = 1e-3 (the differentiation
    workload's shape)."""
    S, N, P, beta, sigma = 8, 512, 360, 1e-3, 0.1
    ang = PH.angles(P).astype(np.float64)
    x = np.stack([PH.random_ellipses(N, 2000 + s) for s in range(S)])
    xs = np.stack([PH.random_ellipses(N, 2100 + s) for s in range(S)])
    from tomocam_amd import gpu
    y = gpu.project(torch_cuda.tensor(xs, dtype=torch_cuda.float32,
                                      device="cuda"), ang).cpu().numpy()
    announce("autograd 8 x 512^2, 360 angles")
    _, g = _loss_grad(torch_cuda, x, y, ang, beta, sigma, torch_cuda.float32)
    x32 = x.astype(np.float32).astype(np.float64)
    data = O.backproject(O.project(x32[:2], ang) - y[:2], ang)
    prior = O.qggmrf_grad_free(x32, sigma, 1.2)[:2]
    assert rel_l2(g[:2], data + beta * prior) < 1e-4


def test_config2_fd_probes(torch_cuda):
    """64 FP64 central finite differences at the differentiation workload's
    shape (8 x 512^2, 360 angles, beta = 1e-3, sigma = 0.1)."""
    from tomocam_amd import autograd as AG
    t = torch_cuda
    S, N, P, beta, sigma, eps = 8, 512, 360, 1e-3, 0.1, 1e-6
    ang = PH.angles(P).astype(np.float64)
    x = np.stack([PH.random_ellipses(N, 2200 + s) for s in range(S)])
    x = x + 0.05 * np.random.default_rng(3).random(x.shape)
    ys = np.stack([PH.random_ellipses(N, 2300 + s) for s in range(S)])
    yt = AG.radon(t.tensor(ys, dtype=t.float64, device="cuda"), ang).detach()
    announce("autograd 8 x 512^2: 64 FP64 finite-difference probes")
    x64 = t.tensor(x, dtype=t.float64, device="cuda")
    grads = {}
    for dt in (t.float32, t.float64):
        xt = x64.to(dt).requires_grad_(True)
        loss = AG.mbir_loss(xt, yt.to(dt), ang, beta=beta, sigma=sigma, p=1.2)
        loss.backward()
        grads[dt] = xt.grad.detach().double()
    gen = t.Generator(device="cuda")
    gen.manual_seed(5)
    fd, got32, got64 = [], [], []
    with t.no_grad():
        for _ in range(64):
            d = t.randn(x.shape, dtype=t.float64, device="cuda", generator=gen)
            d /= t.linalg.norm(d)
            lp = float(AG.mbir_loss(x64 + eps * d, yt, ang, beta=beta,
                                    sigma=sigma, p=1.2))
            lm = float(AG.mbir_loss(x64 - eps * d, yt, ang, beta=beta,
                                    sigma=sigma, p=1.2))
            fd.append((lp - lm) / (2 * eps))
            got32.append(float(t.sum(grads[t.float32] * d)))
            got64.append(float(t.sum(grads[t.float64] * d)))
    fd, got32, got64 = np.array(fd), np.array(got32), np.array(got64)
    scale = float(t.linalg.norm(grads[t.float64]))
    print(f"max |fd - g.d| / |g|: fp32 {np.abs(got32 - fd).max() / scale:.2e}"
          f", fp64 {np.abs(got64 - fd).max() / scale:.2e}")
    assert np.all(np.abs(got32 - fd) <= 1e-4 * scale)
    assert np.all(np.abs(got64 - fd) <= 1e-6 * scale)
