"""Asynchronous operators: tc_project, tc_backproject and tc_gradient leave
their work pending on the caller's stream and take temporaries from the
calling thread's device arena, reused in stream order.  These tests run
calls back to back without synchronising, on one stream and across two
streams, and require every result to equal (bit for bit: the kernels are
deterministic) the same call made alone and synchronised.  The device-output
prior value equals the host one bit for bit."""

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _inputs(torch, k, S=3, N=96):
    g = torch.Generator(device="cuda")
    g.manual_seed(1234 + k)
    return torch.rand((S, N, N), device="cuda", generator=g)


def test_back_to_back_same_stream(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    announce("async projectors, one stream")
    ang = np.arange(60) * np.pi / 60
    xs = [_inputs(torch, k) for k in range(6)]
    torch.cuda.synchronize()
    ref = []
    for x in xs:
        y = gpu.project(x, ang)
        b = gpu.backproject(y, ang)
        torch.cuda.synchronize()
        ref.append((y.clone(), b.clone()))
    out = []
    for x in xs:  # no synchronisation between the calls
        y = gpu.project(x, ang)
        out.append((y, gpu.backproject(y, ang)))
    torch.cuda.synchronize()
    for (y, b), (yr, br) in zip(out, ref):
        assert torch.equal(y, yr)
        assert torch.equal(b, br)


def test_alternating_streams(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    announce("async projectors, two streams")
    ang = np.arange(48) * np.pi / 48
    xs = [_inputs(torch, 10 + k, S=4, N=128) for k in range(6)]
    torch.cuda.synchronize()
    ref = [gpu.backproject(gpu.project(x, ang), ang).clone() for x in xs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    out = []
    for k, x in enumerate(xs):
        s = streams[k % 2]
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            y = gpu.project(x, ang)
            b = gpu.backproject(y, ang)
            out.append(b)
    torch.cuda.synchronize()
    for b, br in zip(out, ref):
        assert torch.equal(b, br)


def test_gradient_async_matches_sync(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    announce("async gradient")
    ang = np.arange(40) * np.pi / 40
    x = _inputs(torch, 20, S=2, N=64)
    b = _inputs(torch, 21, S=2, N=64)
    torch.cuda.synchronize()
    g0 = gpu.gradient(x, b, ang)
    torch.cuda.synchronize()
    g0 = g0.clone()
    gs = [gpu.gradient(x, b, ang) for _ in range(4)]
    torch.cuda.synchronize()
    for g in gs:
        assert torch.equal(g, g0)


def test_prior_value_on_device(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    announce("prior value on the device")
    x = _inputs(torch, 30, S=5, N=70)
    v_host = gpu.qggmrf_value_free(x, 0.3, 1.2)
    v_dev = gpu.qggmrf_value_free_tensor(x, 0.3, 1.2)
    assert v_dev.dtype == torch.float64 and v_dev.dim() == 0
    assert float(v_dev) == v_host
    x64 = x.double()
    assert float(gpu.qggmrf_value_free_tensor(x64, 0.3, 1.2)) == \
        gpu.qggmrf_value_free(x64, 0.3, 1.2)
