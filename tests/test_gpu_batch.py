"""Uneven batching (tc_mbir2_batch): many stacks of different image sizes,
angle sets and slice counts in one call.

Each stack must equal a separate tc_mbir2 call bit for bit (same kernels,
same plans, deterministic reductions), and the FP64 oracle's mbir2 to the
relL2 1e-4 bound.  Errors name the failing stack.
"""

import numpy as np
import pytest
import torch

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu

# (slices, angles, rays)
SHAPES = [(2, 24, 32), (1, 30, 48), (3, 18, 40), (1, 37, 33)]


def _stacks(torch_mod):
    out = []
    for k, (S, P, N) in enumerate(SHAPES):
        ang = PH.angles(P)
        img = np.stack([PH.random_ellipses(N, 300 + 10 * k + s)
                        for s in range(S)])
        sino = O.project(img, ang.astype(np.float64), exact=True)
        out.append((torch_mod.tensor(sino, dtype=torch_mod.float32).cuda(),
                    ang, N / 2.0, sino))
    return out


def test_batch_equals_single_calls_and_oracle(torch_cuda):
    from tomocam_amd import gpu
    st = _stacks(torch_cuda)
    announce("mbir2 batch of 4 uneven stacks")
    recs, infos = gpu.mbir_batch([(s, a, c) for s, a, c, _ in st], 10, 10.0)
    for k, (s, a, c, host) in enumerate(st):
        one, info = gpu.mbir(s, a, c, 10, 10.0)
        assert torch.equal(recs[k], one), f"stack {k}"
        assert infos[k]["padded"] == info["padded"]
    s, a, c, host = st[2]
    ref = O.mbir2(host, a.astype(np.float64), c, 10, 10.0, exact=True)
    assert rel_l2(recs[2].cpu().numpy(), ref) < 1e-4


def test_batch_error_names_the_stack(torch_cuda):
    from tomocam_amd import gpu, _lib
    st = _stacks(torch_cuda)[:2]
    stacks = [(st[0][0], st[0][1], st[0][2]),
              (torch.zeros_like(st[1][0]), st[1][1], st[1][2])]
    announce("mbir2 batch error path")
    with pytest.raises((ValueError, _lib.TomocamError), match="stack 1"):
        gpu.mbir_batch(stacks, 2, 10.0)


def test_empty_batch(torch_cuda):
    from tomocam_amd import gpu
    recs, infos = gpu.mbir_batch([], 5, 10.0)
    assert recs == [] and infos == []


def test_batch_concurrent_streams_equal_sequential(torch_cuda, monkeypatch):
    """Stacks on 4 concurrent streams (the default) against the same batch
    run in order on one stream: identical reconstructions.  Sizes with
    compile-time FFT plans (padded 362 and 726) and one without (padded
    44), largest first in the concurrent schedule."""
    from tomocam_amd import gpu
    stacks = []
    for k, (S, P, N) in enumerate([(1, 60, 256), (3, 90, 512), (2, 24, 32),
                                   (2, 45, 256), (1, 120, 512)]):
        ang = PH.angles(P)
        img = torch.tensor(np.stack([PH.random_ellipses(N, 700 + 10 * k + s)
                                     for s in range(S)]),
                           dtype=torch.float32).cuda()
        stacks.append((gpu.project(img, ang), ang, N / 2.0))
    announce("mbir2 batch sequential (1 stream)")
    monkeypatch.setenv("TOMOCAM_HIP_BATCH_STREAMS", "1")
    seq, infos_seq = gpu.mbir_batch(stacks, 8, 100.0)
    announce("mbir2 batch concurrent (4 streams)")
    monkeypatch.setenv("TOMOCAM_HIP_BATCH_STREAMS", "4")
    con, infos_con = gpu.mbir_batch(stacks, 8, 100.0)
    for k in range(len(stacks)):
        assert infos_con[k]["padded"] == infos_seq[k]["padded"]
        assert infos_con[k]["iterations"] == 8
        # same kernels and group sizes: equal up to FP64 dot order
        assert rel_l2(con[k].cpu().numpy(), seq[k].cpu().numpy()) < 1e-6, k


def test_batch_concurrent_error_names_lowest_stack(torch_cuda, monkeypatch):
    from tomocam_amd import gpu, _lib
    st = _stacks(torch_cuda)
    stacks = [(s, a, c) for s, a, c, _ in st]
    stacks[1] = (torch.zeros_like(stacks[1][0]), stacks[1][1], stacks[1][2])
    stacks[3] = (torch.zeros_like(stacks[3][0]), stacks[3][1], stacks[3][2])
    monkeypatch.setenv("TOMOCAM_HIP_BATCH_STREAMS", "4")
    announce("mbir2 concurrent batch error path")
    with pytest.raises((ValueError, _lib.TomocamError), match="stack [13]"):
        gpu.mbir_batch(stacks, 2, 10.0)


def test_concurrent_host_threads(torch_cuda):
    """Two host threads calling tc_mbir2 at once (each on its own stream)
    get the same results as one after the other: every thread has NUFFT
    plans of its own."""
    import threading
    from tomocam_amd import gpu
    st = _stacks(torch_cuda)
    jobs = [(s, a, c) for s, a, c, _ in st[:2]]
    seq = [gpu.mbir(s, a, c, 6, 10.0)[0] for s, a, c in jobs]
    out = [None, None]
    errs = []

    def run(k):
        try:
            s, a, c = jobs[k]
            strm = torch.cuda.Stream()
            with torch.cuda.stream(strm):
                r, _ = gpu.mbir(s, a, c, 6, 10.0)
            strm.synchronize()
            out[k] = r
        except Exception as e:  # reported below
            errs.append(e)

    announce("two host threads calling mbir2")
    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for k in range(2):
        assert torch.equal(out[k], seq[k]), k
