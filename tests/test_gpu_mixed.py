"""Mixed-precision form on the GPU (through the C-ABI).

The MBIR form stores the gradient in FP16 under a per-slice power-of-two
scale, computes every FFT butterfly and the spreading in FP32 with FP32
spectrum and Toeplitz intermediates, keeps the iterate, T x and b in FP32,
sums the line-search dots in FP64 and recomputes T sol in FP32 every 25
accepted iterations.  The FP16-storage Toeplitz operator (tc_toeplitz_apply
_mixed) also stores its intermediates and the PSF spectrum in FP16.

Tolerances:
* the FP16-storage Toeplitz operator: relative L2 <= 1e-3 against the FP64
  oracle (an FP16 rounding of the input and of each intermediate, each about
  2e-4 RMS);
* MBIR at sigma = 100 (the reference's default smoothness 0.01), 100
  iterations: relative L2 <= 1e-4 against the FP64 oracle of the same
  iteration, and <= 1e-4 against the FP32 form at the 1024^2 / 720-angle
  shape.
"""

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from tomocam_amd import phantom as PH
from conftest import rel_l2, announce

pytestmark = pytest.mark.gpu


def _sino(N, P, S, seed, noise=0.01):
    ang = PH.angles(P)
    img = np.stack([PH.random_ellipses(N, seed + s) for s in range(S)])
    sino = O.project(img, ang.astype(np.float64))
    sino = sino + np.random.default_rng(seed).normal(
        0, noise * sino.max(), sino.shape)
    return ang, sino


@pytest.mark.parametrize("n,S,M", [(340, 2, 750), (362, 2, 768),
                                   (726, 2, 1458), (1024, 1, 2250),
                                   (1090, 1, 2304), (1454, 2, 2916),
                                   (2048, 1, 4500), (2180, 1, 4608)])
def test_toeplitz_mixed_sizes(torch_cuda, n, S, M):
    """FFT lengths 750, 768, 1458, 2250, 2304, 2916, 4500, 4608 (the
    polar-grid PSF up to 362, a synthetic even PSF above): random and
    piecewise-constant inputs, and an all-zero slice (scale 1, exact zeros
    out)."""
    torch = torch_cuda
    from tomocam_amd import gpu
    ang = PH.angles(180)
    ps = O.psf(ang, n) if n <= 362 else None
    if ps is None:
        r = np.random.default_rng(n).standard_normal((2 * n - 1, 2 * n - 1))
        ps = 0.5 * (r + r[::-1, ::-1])
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    assert psf.fft_size == M
    rng = np.random.default_rng(n + 3)
    for kind in ("random", "ellipses"):
        if kind == "random":
            x = rng.standard_normal((S, n, n))
        else:
            x = np.stack([PH.random_ellipses(n, 40 + s) for s in range(S)])
        ref = O.toeplitz(x, ps)
        announce(f"toeplitz mixed n={n} S={S} {kind}")
        out, xtx = psf.convolve_mixed(
            torch.tensor(x, dtype=torch.float32).cuda(), with_dot=True)
        out = out.cpu().numpy()
        assert np.all(np.isfinite(out))
        assert rel_l2(out, ref) < 1e-3
        assert abs(xtx - np.sum(x * ref)) <= 1e-3 * np.sum(np.abs(x * ref))
    if S > 1:
        x = rng.standard_normal((S, n, n))
        x[1] = 0.0
        announce(f"toeplitz mixed n={n} zero slice")
        out = psf.convolve_mixed(torch.tensor(x, dtype=torch.float32).cuda())
        out = out.cpu().numpy()
        assert np.all(out[1] == 0.0)
        assert rel_l2(out[0], O.toeplitz(x[:1], ps)[0]) < 1e-3


def test_toeplitz_mixed_needs_compile_time_plan(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    n = 181  # FFT length 378: generic kernels only
    ps = O.psf(PH.angles(30), n)
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    announce("toeplitz mixed without plan")
    with pytest.raises(ValueError, match="compile-time FFT plan"):
        psf.convolve_mixed(torch.zeros((1, n, n), device="cuda"))


def test_mbir2_mixed_vs_oracle(torch_cuda):
    """4 x 256^2 (padded 362, FFT length 768), 180 angles, 100 iterations,
    sigma = 100: both precisions against the FP64 oracle."""
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P, S, iters, sigma = 256, 180, 2, 100, 100.0
    ang, sino = _sino(N, P, S, 11)
    a64 = ang.astype(np.float64)
    st = O.mbir2_setup(sino, a64, N / 2)
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, sigma, 1.2)
    rec = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, iters, st.step)
    ref = O.postproc(rec, N)
    st_t = torch.tensor(sino, dtype=torch.float32).cuda()
    announce("mbir2 mixed vs oracle 2x256^2")
    mixed, info = gpu.mbir(st_t, ang, N / 2, iters, sigma, precision="mixed")
    plain, _ = gpu.mbir(st_t, ang, N / 2, iters, sigma)
    e_mixed = rel_l2(mixed.cpu().numpy(), ref)
    e_plain = rel_l2(plain.cpu().numpy(), ref)
    print(f"relL2 vs FP64: mixed {e_mixed:.3e}  fp32 {e_plain:.3e}")
    assert info["iterations"] == iters
    assert info["saturated"] == 0
    assert info["padded"] == 362
    assert e_plain < 1e-4
    assert e_mixed < 1e-4


def test_mixed_slice_groups_and_streams(torch_cuda, monkeypatch):
    """The mixed-precision MBIR (FP16 gradient, FP32 intermediates) and the
    FP16-intermediate operator (per-group pass-B scales) on 5 slices in
    groups of 1 or 2, on one stream and round-robin over two or three,
    against one group on one stream."""
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P, S, iters, sigma = 256, 120, 5, 20, 100.0
    ang, sino = _sino(N, P, S, 23)
    st_t = torch.tensor(sino, dtype=torch.float32).cuda()
    ps = O.psf(PH.angles(180), 362)
    psf = gpu.PointSpreadFunction(
        psf_array=torch.tensor(ps, dtype=torch.float32).cuda())
    x = torch.tensor(np.random.default_rng(5).standard_normal((S, 362, 362)),
                     dtype=torch.float32).cuda()
    monkeypatch.setenv("TOMOCAM_HIP_AUTOTUNE", "0")
    recs, convs = {}, {}
    for groups, streams in (("0", "1"), ("2", "2"), ("1", "3"), ("2", "3")):
        monkeypatch.setenv("TOMOCAM_HIP_TOEPLITZ_SLICES", groups)
        monkeypatch.setenv("TOMOCAM_HIP_TOEPLITZ_STREAMS", streams)
        announce(f"mixed slice groups {groups} streams {streams}")
        rec, info = gpu.mbir(st_t, ang, N / 2, iters, sigma,
                             precision="mixed")
        assert info["saturated"] == 0
        recs[(groups, streams)] = rec.cpu().numpy()
        convs[(groups, streams)] = psf.convolve_mixed(x).cpu().numpy()
    one, cone = recs[("0", "1")], convs[("0", "1")]
    for key in (("2", "2"), ("1", "3"), ("2", "3")):
        assert rel_l2(recs[key], one) < 1e-6, key
        assert rel_l2(convs[key], cone) < 1e-6, key


def test_mbir2_mixed_power_lipschitz(torch_cuda):
    """L from 20 power-iteration steps on T (FP64 dots) against the FP64
    oracle's 20 steps; the step is min(1, 1/L)."""
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P, S = 256, 180, 1
    ang, sino = _sino(N, P, S, 12)
    st = O.mbir2_setup(sino, ang.astype(np.float64), N / 2)
    lam = O.power_lipschitz(st.psf, st.x0.shape[-1], 20)
    announce("mbir2 mixed power lipschitz")
    _, info = gpu.mbir(torch.tensor(sino, dtype=torch.float32).cuda(), ang,
                       N / 2, 2, 100.0, precision="mixed", lipschitz_iters=20)
    assert abs(info["lipschitz"] - lam) <= 1e-4 * lam
    assert info["step"] == pytest.approx(min(1.0, 1.0 / info["lipschitz"]),
                                         rel=1e-12)


def test_mbir2_mixed_vs_fp32_1024_shape(torch_cuda):
    """4 slices of the 32 x 1024^2 / 720-angle workload (padded 1454, FFT
    length 2916), 30 iterations, sigma = 100: mixed against FP32."""
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P, S, iters, sigma = 1024, 720, 4, 30, 100.0
    ang = PH.angles(P)
    img = torch.tensor(np.stack([PH.random_ellipses(N, 1000 + s)
                                 for s in range(S)]),
                       dtype=torch.float32).cuda()
    announce("mbir2 mixed vs fp32 4x1024^2")
    sino = gpu.project(img, ang)
    sino = sino + 0.01 * sino.max() * torch.randn(
        sino.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    mixed, info = gpu.mbir(sino, ang, N / 2, iters, sigma, precision="mixed")
    plain, _ = gpu.mbir(sino, ang, N / 2, iters, sigma)
    e = rel_l2(mixed.cpu().numpy(), plain.cpu().numpy())
    print(f"relL2 mixed vs fp32: {e:.3e}")
    assert info["padded"] == 1454
    assert info["iterations"] == iters
    assert info["saturated"] == 0
    assert e < 1e-4


def test_mbir2_mixed_divergent_sigma_stays_finite(torch_cuda):
    """sigma = 0.1 (the benchmark's divergent setting): the FP16 gradient's
    scale follows the growth; the result stays finite."""
    torch = torch_cuda
    from tomocam_amd import gpu
    N, P, S, iters = 256, 180, 2, 20
    ang, sino = _sino(N, P, S, 13)
    announce("mbir2 mixed sigma 0.1")
    out, info = gpu.mbir(torch.tensor(sino, dtype=torch.float32).cuda(), ang,
                         N / 2, iters, 0.1, precision="mixed")
    assert info["iterations"] == iters
    assert np.all(np.isfinite(out.cpu().numpy()))


def test_mbir2_mixed_argument_checks(torch_cuda):
    torch = torch_cuda
    from tomocam_amd import gpu
    sino = torch.ones((1, 30, 100), device="cuda")
    ang = PH.angles(30)
    announce("mbir2 mixed argument checks")
    # width 100 pads to 142: FFT length 288 has no compile-time plan
    with pytest.raises(ValueError, match="compile-time FFT plan"):
        gpu.mbir(sino, ang, 50.0, 2, 100.0, precision="mixed")
    with pytest.raises(ValueError):
        gpu.mbir(sino, ang, 50.0, 2, 100.0, precision="bf16")
    with pytest.raises(ValueError):
        gpu.mbir(sino, ang, 50.0, 2, 100.0, lipschitz_iters=5)


def test_python_entry_points_precision_keyword(torch_cuda):
    """tomocam.MBIR / recon / cTomocam.mbir take precision="mixed" as a
    keyword (the positional reference signature is unchanged)."""
    import tomocam_amd as tomocam
    from tomocam_amd import cTomocam
    N, P, S = 256, 90, 1
    ang, sino = _sino(N, P, S, 14)
    s32 = sino.astype(np.float32)
    announce("MBIR precision keyword")
    plain = tomocam.MBIR(s32, ang, N / 2, 20, 0.01)
    mixed = tomocam.MBIR(s32, ang, N / 2, 20, 0.01, precision="mixed")
    assert rel_l2(mixed, plain) < 1e-4
    again = tomocam.recon(s32, ang, N / 2, num_iters=20, precision="mixed")
    assert np.array_equal(again, mixed)
    direct = cTomocam.mbir(s32, ang, N / 2, 20, 100.0, 1e-4, 1e-4,
                           precision="mixed")
    assert np.array_equal(direct, mixed)
    with pytest.raises(ValueError):
        tomocam.MBIR(s32, ang, N / 2, 2, 0.01, precision="fp8")
