"""HIP-graph replay of the steady-state MBIR attempt against eager launches.

With TOMOCAM_HIP_GRAPHS=1 the MBIR loop captures one attempt (stencil,
Toeplitz passes with the fused commit, dot reduction, copy of the dots to the
host) per buffer role and replays it; by default (or =0) every kernel is
launched eagerly.  Both run
the same kernels in the same order on the same data, so the results, every
iteration's error and x-error, and the attempt counts must be bit-identical:
each replay is checked through the error it feeds into the host's accept
test, with the host's own synchronisation and eager work (rejected
attempts, FP32 refreshes) between replays.  The eager path itself is pinned
against the FP64 oracle by the other MBIR tests.
"""

import os

import numpy as np
import pytest

from tomocam_amd import phantom as PH
from conftest import announce

pytestmark = pytest.mark.gpu

N, P, S = 256, 180, 3   # padded width 362, FFT length 768 (compile-time passes)


def _timers_off(gpu):
    for k in gpu.KERNEL_CLASSES:
        gpu.ktimer_enable(k, False)


def _with_graphs(on, f):
    old = os.environ.get("TOMOCAM_HIP_GRAPHS")
    os.environ["TOMOCAM_HIP_GRAPHS"] = "1" if on else "0"
    try:
        return f()
    finally:
        if old is None:
            del os.environ["TOMOCAM_HIP_GRAPHS"]
        else:
            os.environ["TOMOCAM_HIP_GRAPHS"] = old


@pytest.fixture(scope="module")
def problem(torch_cuda):
    """x_true, b = T x_true and ||y||^2 = x_true . T x_true on the PSF of a
    180-angle geometry at width 362, so err(x)^2 = (x - x_true).T(x - x_true)."""
    torch = torch_cuda
    from tomocam_amd import gpu
    psf = gpu.PointSpreadFunction(PH.angles(P), 362)
    assert psf.fft_size == 768
    xt = np.stack([PH.random_ellipses(362, 71 + s) for s in range(S)])
    xt = torch.tensor(xt, dtype=torch.float32).cuda()
    b = psf.convolve(xt)
    ynorm = float((xt.double() * b.double()).sum())
    lip = float(psf.convolve(torch.ones_like(xt)).max())
    return psf, xt, b, ynorm, min(1.0, 1.0 / lip)


def _compare(torch, gpu, run):
    eager = _with_graphs(False, run)
    before = gpu.graph_replays()
    graph = _with_graphs(True, run)
    replays = gpu.graph_replays() - before
    assert torch.equal(eager[0], graph[0])
    np.testing.assert_array_equal(eager[1], graph[1])
    np.testing.assert_array_equal(eager[2], graph[2])
    assert eager[3] == graph[3]
    steady = eager[3]["attempts"] - eager[3]["backtracks"]
    # per buffer role the first attempt is eager, the second captures
    assert replays >= steady - 8, (replays, eager[3])
    return eager[3]


@pytest.mark.parametrize("refresh", [10, 3])
def test_run2_graph_replay_bit_identical(torch_cuda, problem, refresh):
    """Steady state; with refresh 3 the FP32 refreshes move T sol to the
    other buffer often, so more buffer roles are captured."""
    torch = torch_cuda
    from tomocam_amd import gpu
    psf, xt, b, ynorm, step = problem
    _timers_off(gpu)
    announce(f"run2 graphs vs eager refresh={refresh}")
    _compare(torch, gpu, lambda: gpu.run2(psf, 0.5 * b, b, 25, step, 100.0,
                                          1.2, ynorm, refresh_every=refresh))


def test_run2_graph_replay_with_rejections(torch_cuda, problem):
    """Rejected attempts (y = sol + beta (sol - x), reduced step) run eagerly
    between replays.  A smooth offset of a nearly converged point with a
    step of about 20 err(x0) / L makes the accept test reject (the FP64
    oracle rejects at 10 and 26 times err / L on this geometry)."""
    torch = torch_cuda
    from tomocam_amd import gpu
    psf, xt, b, ynorm, step = problem
    _timers_off(gpu)
    sigma = 1e4
    announce("run2 graphs vs eager with rejections: converging 300 iterations")
    xs = _with_graphs(False, lambda: gpu.run2(psf, 0.5 * b, b, 300, step,
                                              sigma, 1.2, ynorm))[0]
    x0 = xs + 1e-3
    tx = psf.convolve(x0)
    fy2 = ynorm + float((x0.double() * tx.double()).sum()) \
        - 2.0 * float((x0.double() * b.double()).sum())
    lip = 1.0 / step
    big = 20.0 * np.sqrt(max(fy2, 0.0)) / lip
    announce(f"run2 graphs vs eager with rejections: step {big:.3g}")
    st = _compare(torch, gpu, lambda: gpu.run2(psf, x0, b, 25, big, sigma,
                                               1.2, ynorm))
    assert st["backtracks"] > 0, st


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_mbir_graph_replay_bit_identical(torch_cuda, precision):
    torch = torch_cuda
    from tomocam_amd import gpu
    _timers_off(gpu)
    ang = PH.angles(P)
    img = np.stack([PH.random_ellipses(N, 91 + s) for s in range(S)])
    sino = gpu.project(torch.tensor(img, dtype=torch.float32).cuda(), ang)
    announce(f"mbir graphs vs eager ({precision})")

    def run():
        return gpu.mbir(sino, ang, N / 2, 30, 100.0, precision=precision)

    eager = _with_graphs(False, run)
    before = gpu.graph_replays()
    graph = _with_graphs(True, run)
    assert gpu.graph_replays() - before >= 20
    assert torch.equal(eager[0], graph[0])
    assert eager[1] == graph[1]


def test_graphs_off_by_default(torch_cuda, problem):
    torch = torch_cuda
    from tomocam_amd import gpu
    psf, xt, b, ynorm, step = problem
    _timers_off(gpu)
    announce("run2 default: no graph replays")
    old = os.environ.pop("TOMOCAM_HIP_GRAPHS", None)
    try:
        before = gpu.graph_replays()
        gpu.run2(psf, 0.5 * b, b, 10, step, 100.0, 1.2, ynorm)
        assert gpu.graph_replays() == before
    finally:
        if old is not None:
            os.environ["TOMOCAM_HIP_GRAPHS"] = old


def test_graphs_off_with_kernel_timing(torch_cuda, problem):
    """Kernel timing records events per launch, so it runs eagerly."""
    torch = torch_cuda
    from tomocam_amd import gpu
    psf, xt, b, ynorm, step = problem
    announce("run2 with kernel timing: no graph replays")
    gpu.ktimer_enable(1, True)
    try:
        before = gpu.graph_replays()
        _with_graphs(True, lambda: gpu.run2(psf, 0.5 * b, b, 10, step, 100.0,
                                            1.2, ynorm))
        assert gpu.graph_replays() == before
        ms, cnt, _ = gpu.ktimer_collect(1)
        assert cnt > 0
    finally:
        _timers_off(gpu)


def test_kernel_timer_stride(torch_cuda, problem):
    """A stride k times one launch in k of a class: the counts of two
    identical runs are n and ceil(n / k), with the same bytes per launch."""
    from tomocam_amd import gpu
    psf, xt, b, ynorm, step = problem
    announce("run2 with kernel timing, stride 1 then 3")
    _timers_off(gpu)
    res = []
    try:
        for stride in (1, 3):
            gpu.ktimer_stride(stride)
            gpu.ktimer_enable(1, True)
            gpu.run2(psf, 0.5 * b, b, 10, step, 100.0, 1.2, ynorm)
            ms, cnt, nb = gpu.ktimer_collect(1)
            gpu.ktimer_enable(1, False)
            res.append((ms, cnt, nb))
    finally:
        gpu.ktimer_stride(1)
        _timers_off(gpu)
    (ms1, n1, b1), (ms3, n3, b3) = res
    assert n1 > 0 and n3 == -(-n1 // 3)
    assert ms3 > 0 and b3 / n3 == pytest.approx(b1 / n1, rel=1e-12)


def test_concurrent_batch_with_graphs_requested(torch_cuda):
    """Concurrent batch workers keep graphs off even when requested (a
    capture in one worker would break the others' synchronous copies) and
    give the sequential results."""
    torch = torch_cuda
    from tomocam_amd import gpu
    _timers_off(gpu)
    stacks = []
    for k, (n, p, s) in enumerate([(256, 120, 2), (128, 90, 3), (256, 150, 1)]):
        ang = PH.angles(p)
        img = np.stack([PH.random_ellipses(n, 300 + 10 * k + i)
                        for i in range(s)])
        sino = gpu.project(torch.tensor(img, dtype=torch.float32).cuda(), ang)
        stacks.append((sino, ang, n / 2))
    announce("mbir_batch, 3 stacks, graphs requested")
    old = os.environ.get("TOMOCAM_HIP_BATCH_STREAMS")
    try:
        os.environ["TOMOCAM_HIP_BATCH_STREAMS"] = "3"
        before = gpu.graph_replays()
        conc, _ = _with_graphs(True, lambda: gpu.mbir_batch(stacks, 20))
        assert gpu.graph_replays() == before
        os.environ["TOMOCAM_HIP_BATCH_STREAMS"] = "1"
        seq, _ = _with_graphs(False, lambda: gpu.mbir_batch(stacks, 20))
    finally:
        if old is None:
            os.environ.pop("TOMOCAM_HIP_BATCH_STREAMS", None)
        else:
            os.environ["TOMOCAM_HIP_BATCH_STREAMS"] = old
    for a, b in zip(conc, seq):
        assert torch.equal(a, b)
