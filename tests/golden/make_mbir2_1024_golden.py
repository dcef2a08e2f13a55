"""Writes mbir2_1024_slice.npz: the FP64 oracle's mbir2 on one slice of the
32 x 1024^2 / 720-angle workload (padded width 1454), 30 iterations,
sigma = 100, sampled on a 64 x 64 grid of pixels (every 16th row and
column) plus whole-image sums.  The oracle's PSF at this size takes minutes,
so the GPU test compares against this fixture instead of running it.

Input (rebuilt identically by the test): random_ellipses(1024, seed 1000),
projected by the oracle's ES-kernel NUFFT projector, plus Gaussian noise of
sd 1 % of the maximum from numpy default_rng(1007).

    python tests/golden/make_mbir2_1024_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import tomocam_oracle as O  # noqa: E402
from tomocam_amd import phantom as PH  # noqa: E402

N, P, ITERS, SIGMA = 1024, 720, 30, 100.0


def make_sinogram():
    ang = PH.angles(P).astype(np.float64)
    img = PH.random_ellipses(N, 1000)[None]
    sino = O.project(img, ang)
    sino = sino + np.random.default_rng(1007).normal(0, 0.01 * sino.max(),
                                                     sino.shape)
    return ang, sino


def main():
    ang, sino = make_sinogram()
    st = O.mbir2_setup(sino, ang, N / 2)
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, SIGMA, 1.2)
    rec = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, ITERS, st.step)
    out = O.postproc(rec, N)[0]
    np.savez(os.path.join(HERE, "mbir2_1024_slice.npz"),
             sampled=out[::16, ::16], l2=np.linalg.norm(out),
             total=out.sum(), lipschitz=st.lipschitz,
             sino_norm=st.sino_norm, step=st.step,
             sino_checksum=np.array([sino.sum(), np.linalg.norm(sino)]))


if __name__ == "__main__":
    main()
