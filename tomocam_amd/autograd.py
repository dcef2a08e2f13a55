"""Differentiable projector and prior on the MI355X kernels (torch.autograd).

The reference has no differentiable form; this module provides it:

- :class:`RadonFunction`: forward is the NUFFT projector P x (the C-ABI
  ``tc_project``); the hand-written backward is its exact adjoint, the
  backprojection kernel P^T g (``tc_backproject``).  Only the angles are
  kept between the two: the oversampled grids are recomputed in the backward
  call, never stored.
- :class:`QGGMRFFunction`: the free-boundary qGGMRF prior R(x) (sum over
  unordered in-bounds neighbour pairs of w_d rho(x_v - x_{v+d}), the
  reference's weights, q = 2, c = 0.001) and its exact gradient
  (``tc_qggmrf_value_free`` / ``tc_qggmrf_grad_free``).  In the interior the
  gradient is the reference stencil (``add_total_var2``); at the faces pairs
  leaving the volume are dropped, so the gradient is exact.

Arithmetic is FP32 on the GPU (value sums in FP64) for float32 and
lower-precision inputs, which are cast to float32; float64 inputs run the
FP64 kernels (``tc_project_f64``, ``tc_backproject_f64``,
``tc_qggmrf_*_free_f64``) end to end.  Gradients come back in the input's
dtype.
"""

import torch

from . import gpu


def _work(t: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """t detached as the arithmetic dtype: float64 stays float64, anything
    else becomes float32."""
    if not t.is_cuda:
        raise ValueError("tomocam_amd.autograd works on CUDA (HIP) tensors")
    wt = torch.float64 if dtype == torch.float64 else torch.float32
    return t.detach().to(wt).contiguous()


class RadonFunction(torch.autograd.Function):
    """y = P x for images [S, N, N] (or [N, N]) and angles [P]."""

    @staticmethod
    def forward(ctx, img: torch.Tensor, angles) -> torch.Tensor:
        # kept in FP64; project/backproject round them for the FP32 kernels
        ang = gpu._angles64(angles)
        ctx.angles = ang
        ctx.dtype = img.dtype
        ctx.two_d = img.dim() == 2
        out = gpu.project(_work(img, img.dtype), ang)
        if ctx.two_d:
            out = out[0]
        return out.to(img.dtype)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        g = _work(grad, ctx.dtype)
        if ctx.two_d:
            g = g[None]
        img_grad = gpu.backproject(g, ctx.angles, fbp=False)
        if ctx.two_d:
            img_grad = img_grad[0]
        return img_grad.to(ctx.dtype), None


class QGGMRFFunction(torch.autograd.Function):
    """R(x) for a volume [S, R, C] (or [R, C]); returns a 0-d tensor."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, sigma: float, p: float) -> torch.Tensor:
        if sigma <= 0:
            raise ValueError("sigma must be > 0")
        xw = _work(x, x.dtype)
        ctx.save_for_backward(xw)
        ctx.sigma, ctx.p, ctx.dtype = float(sigma), float(p), x.dtype
        # the value stays on the device: no host synchronisation per step
        return gpu.qggmrf_value_free_tensor(xw, ctx.sigma, ctx.p).to(x.dtype)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        (xw,) = ctx.saved_tensors
        g = torch.zeros_like(xw)
        gpu.qggmrf_grad_free(xw, g, ctx.sigma, ctx.p)
        return (g * grad.to(xw.dtype)).to(ctx.dtype), None, None


def radon(img: torch.Tensor, angles) -> torch.Tensor:
    """Differentiable Radon transform (backward = backprojection)."""
    return RadonFunction.apply(img, angles)


def qggmrf(x: torch.Tensor, sigma: float, p: float = 1.2) -> torch.Tensor:
    """Differentiable free-boundary qGGMRF prior value."""
    return QGGMRFFunction.apply(x, sigma, p)


def mbir_loss(x: torch.Tensor, sino: torch.Tensor, angles, beta: float = 1e-3,
              sigma: float = 0.1, p: float = 1.2) -> torch.Tensor:
    """0.5 ||P x - y||^2 + beta R(x), differentiable in x."""
    r = radon(x, angles) - sino
    loss = 0.5 * torch.sum(r * r)
    if beta:
        loss = loss + beta * qggmrf(x, sigma, p)
    return loss
