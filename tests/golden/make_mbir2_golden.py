"""Writes mbir2_<N>_slice.npz: the FP64 oracle's mbir2 on one slice of a
workload shape, 30 iterations, sigma = 100, sampled on a 64 x 64 grid of
pixels plus whole-image sums and the set-up scalars; and
mbir2_2048_slice_100it.npz, the same for the 2048 shape at the workload's
100 iterations (the benchmark's accuracy check against FP64).  The oracle's PSF at
these sizes takes minutes (1024) to tens of minutes (2048), so the GPU tests
compare against these fixtures instead of running it.

Shapes (one slice each):
  1024: the 32 x 1024^2 / 720-angle workload (padded width 1454);
  2048: the 64 x 2048^2 / 1501-angle workload (padded width 2908).
Input (rebuilt identically by the tests): random_ellipses(N, seed), projected
by the oracle's ES-kernel NUFFT projector, plus Gaussian noise from numpy
default_rng(noise seed) with sd a fraction of the maximum.

    python tests/golden/make_mbir2_golden.py 1024
    python tests/golden/make_mbir2_golden.py 2048
    python tests/golden/make_mbir2_golden.py 2048 100
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import tomocam_oracle as O  # noqa: E402
from tomocam_amd import phantom as PH  # noqa: E402

ITERS, SIGMA = 30, 100.0
# N: (projections, phantom seed, noise fraction, noise seed)
SHAPES = {1024: (720, 1000, 0.01, 1007), 2048: (1501, 3000, 0.005, 3007)}


def fixture_path(N, iters=ITERS):
    tag = "" if iters == ITERS else f"_{iters}it"
    return os.path.join(HERE, f"mbir2_{N}_slice{tag}.npz")


def make_sinogram(N):
    P, seed, frac, nseed = SHAPES[N]
    ang = PH.angles(P).astype(np.float64)
    img = PH.random_ellipses(N, seed)[None]
    sino = O.project(img, ang)
    sino = sino + np.random.default_rng(nseed).normal(0, frac * sino.max(),
                                                      sino.shape)
    return ang, sino


def main(N, iters=ITERS):
    ang, sino = make_sinogram(N)
    st = O.mbir2_setup(sino, ang, N / 2)
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, SIGMA, 1.2)
    rec = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, iters, st.step)
    out = O.postproc(rec, N)[0]
    k = N // 64
    np.savez(fixture_path(N, iters), sampled=out[::k, ::k], stride=k,
             iterations=iters,
             l2=np.linalg.norm(out), total=out.sum(), lipschitz=st.lipschitz,
             sino_norm=st.sino_norm, step=st.step,
             sino_checksum=np.array([sino.sum(), np.linalg.norm(sino)]))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1024,
         int(sys.argv[2]) if len(sys.argv) > 2 else ITERS)
