"""qGGMRF stencil on the GPU vs the oracle restatement of test/tvd.cpp.

Inputs follow test/tvd.cpp:80-99: NPRandom(5489) uniform [0,1) values,
sigma = 10, p = 1.2 (plus sigma = 0.1 and 100).  Tolerance: relative L2
<= 1e-5 and max abs error <= 1e-5 * max|g| against the FP64 restatement
(rho' is evaluated in FP32 on the GPU).
"""

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


def _run(torch, x, sigma, p, g0=None, fn="add_total_var2"):
    from tomocam_amd import gpu
    xt = torch.tensor(x, dtype=torch.float32).cuda()
    g = torch.zeros_like(xt) if g0 is None else torch.tensor(
        g0, dtype=torch.float32).cuda()
    getattr(gpu, fn)(xt, g, sigma, p)
    torch.cuda.synchronize()
    return g.cpu().numpy()


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 5, 7), (2, 17, 65),
                                   (3, 64, 64), (5, 70, 130), (16, 127, 255)])
@pytest.mark.parametrize("sigma", [10.0, 0.1])
def test_qggmrf_vs_oracle(torch_cuda, shape, sigma):
    rs = O.NPRandom()
    x = rs.rand(int(np.prod(shape))).reshape(shape)
    ref = O.qggmrf_grad(x.astype(np.float32), sigma, 1.2)
    announce(f"qggmrf shape={shape} sigma={sigma}")
    got = _run(torch_cuda, x, sigma, 1.2)
    if np.abs(ref).max() == 0:
        assert np.abs(got).max() == 0
        return
    assert rel_l2(got, ref) < 1e-5
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_qggmrf_matches_tvd_cpp_loops(torch_cuda):
    """Loop-for-loop restatement of test/tvd.cpp:39-75 in float32."""
    shape = (3, 9, 11)
    x = O.NPRandom().rand(int(np.prod(shape))).reshape(shape).astype(np.float32)
    ref = O.qggmrf_grad_loops(x, 10.0, 1.2)
    announce("qggmrf tvd.cpp loops")
    got = _run(torch_cuda, x, 10.0, 1.2)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_qggmrf_accumulates_and_variant(torch_cuda):
    """g is incremented in place; add_total_var gives the same arithmetic."""
    rng = np.random.default_rng(5)
    x = rng.random((4, 33, 40))
    g0 = rng.standard_normal(x.shape)
    ref = g0 + O.qggmrf_grad(x.astype(np.float32), 100.0, 1.2)
    announce("qggmrf accumulate")
    got = _run(torch_cuda, x, 100.0, 1.2, g0=g0)
    assert rel_l2(got, ref) < 1e-5
    got2 = _run(torch_cuda, x, 100.0, 1.2, g0=g0, fn="add_total_var")
    assert np.array_equal(got, got2)


def test_qggmrf_large_tvd_shape(torch_cuda):
    """The tvd.cpp volume width (1023) with 4 slices."""
    shape = (4, 1023, 1023)
    x = O.NPRandom().rand(int(np.prod(shape))).reshape(shape)
    x32 = x.astype(np.float32)
    announce("qggmrf 4x1023x1023")
    got = _run(torch_cuda, x32, 10.0, 1.2)
    ref = O.qggmrf_grad(x32, 10.0, 1.2)
    assert rel_l2(got, ref) < 1e-5
