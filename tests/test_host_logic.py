"""Host-side logic without a GPU: the Toeplitz FFT length rule (through the
C-ABI) and the benchmark's VALU-roofline arithmetic."""

import ctypes as C
import math

import pytest


@pytest.mark.parametrize("n,M", [(362, 768), (340, 750), (726, 1458),
                                 (1090, 2304),
                                 (1454, 2916), (2180, 4608), (2908, 5832),
                                 (2048, 4500), (2250, 4608), (2304, 4608),
                                 (1024, 2250), (1152, 2304),
                                 (181, 378), (44, 90), (2, 4)])
def test_toeplitz_fft_length(n, M):
    """M >= 2n - 1; the compile-time plan sizes for the workload widths; where
    both 4500 and 4608 (2250 and 2304, 750 and 768) fit, the larger (its plan
    is 1.54x (1.17x, 1.22x) cheaper per point); the smallest even
    2,3,5,7-smooth length where no plan is within 10 %."""
    from tomocam_amd import _lib
    m = C.c_int()
    _lib.call("tc_toeplitz_fft_length", n, C.byref(m))
    assert m.value == M and m.value >= 2 * n - 1
    with pytest.raises(ValueError):
        _lib.call("tc_toeplitz_fft_length", 1, C.byref(m))


def test_valu_roofline_arithmetic():
    import bench
    n, L = 1454, 2916
    H = L // 2 + 1
    for wtb, specb, ns in ((8, 4, 8), (4, 2, 16), (8, 2, 4)):
        # slices x (Wt in + out) + the spectrum, per launch
        per = ns * 2 * wtb * n * H + specb * H * L
        r = bench.valu_roofline({"bytes_per_launch": per, "avg_ms": 1.0},
                                "toeplitz_cols", n, L, wtb, specb)
        assert r["slices_per_launch"] == pytest.approx(ns)
        flops = ns * H * (10 * L * math.log2(L) + 2 * L)
        assert r["flops_per_launch"] == pytest.approx(flops)
        assert r["achieved"] == pytest.approx(flops / 1e-3 / 1e12)
