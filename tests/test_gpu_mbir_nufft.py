"""tomocam::mbir (src/mbir.cpp:38-119), the NUFFT normal-operator variant,
against the FP64 oracle restatement (exact NUDFT at these sizes):
relL2 <= 1e-4 on the reconstruction and the same accepted/backtracked
counts; the degenerate sinogram (max == min) returns the guess unchanged.
"""

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,P,S,center,sigma", [(32, 24, 2, 16.0, 10.0),
                                                (33, 30, 1, 15.3, 100.0)])
def test_mbir_nufft_matches_oracle(torch_cuda, N, P, S, center, sigma):
    from tomocam_amd import gpu
    ang = PH.angles(P)
    img = np.stack([PH.random_ellipses(N, 500 + s) for s in range(S)])
    sino = O.project(img, ang.astype(np.float64), exact=True)
    x0 = O.backproject(sino / sino.max(), ang.astype(np.float64), fbp=True,
                       exact=True)
    iters = 8
    log = O.Run2Log()
    ref = O.mbir_nufft(x0, sino, ang.astype(np.float64), center, iters, sigma,
                       exact=True, log=log)
    announce(f"mbir (NUFFT variant) N={N} P={P} S={S}")
    t = torch_cuda
    out, info = gpu.mbir_nufft(t.tensor(sino, dtype=t.float32).cuda(), ang,
                               center, t.tensor(x0, dtype=t.float32).cuda(),
                               iters, sigma)
    assert int(info["iterations"]) == iters
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4


def test_mbir_nufft_constant_sinogram_returns_guess(torch_cuda):
    from tomocam_amd import gpu
    t = torch_cuda
    sino = t.full((1, 12, 16), 3.0, device="cuda")
    x0 = t.rand(1, 16, 16, device="cuda")
    announce("mbir (NUFFT variant) constant sinogram")
    out, info = gpu.mbir_nufft(sino, PH.angles(12), 8.0, x0, 5, 10.0)
    assert t.equal(out, x0)
    assert info["padded"] == 0
