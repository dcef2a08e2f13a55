"""MBIR at the 1024^2 / 720-angle workload shape (padded width 1454, FFT
length 2916), 30 iterations, against the FP64 oracle, through a committed
fixture: the
oracle's PSF at this size takes minutes, so tests/golden/
make_mbir2_1024_golden.py ran it once and stored the result sampled on a
64 x 64 pixel grid plus whole-image sums.  Tolerance: relative 1e-4 (FP32
and mixed precision), 1e-5 on the set-up scalars."""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2, announce

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_mbir2_1024_golden as G  # noqa: E402

FIX = os.path.join(ROOT, "tests", "golden", "mbir2_1024_slice.npz")


@pytest.fixture(scope="module")
def case():
    ang, sino = G.make_sinogram()
    gold = np.load(FIX)
    # the same input as the fixture's (numpy on the host)
    chk = np.array([sino.sum(), np.linalg.norm(sino)])
    np.testing.assert_allclose(chk, gold["sino_checksum"], rtol=1e-12)
    return ang, sino, gold


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_mbir2_1024_slice_vs_oracle(torch_cuda, case, precision):
    torch = torch_cuda
    from tomocam_amd import gpu
    ang, sino, gold = case
    announce(f"mbir2 1024^2 slice vs oracle fixture ({precision})")
    out, info = gpu.mbir(torch.tensor(sino, dtype=torch.float32).cuda(), ang,
                         G.N / 2, G.ITERS, G.SIGMA, precision=precision)
    img = out.cpu().numpy()[0].astype(np.float64)
    assert info["padded"] == 1454
    assert abs(info["lipschitz"] - gold["lipschitz"]) <= 1e-5 * gold["lipschitz"]
    assert abs(info["sino_norm"] - gold["sino_norm"]) <= 1e-5 * gold["sino_norm"]
    assert info["step"] == gold["step"]
    e = rel_l2(img[::16, ::16], gold["sampled"])
    print(f"relL2 on the sampled grid ({precision}): {e:.3e}")
    assert e < 1e-4
    assert abs(np.linalg.norm(img) - gold["l2"]) <= 1e-4 * gold["l2"]
    assert abs(img.sum() - gold["total"]) <= 1e-4 * abs(gold["total"])
