"""The MFMA decision measurement: tests/cpp/mfma_stage runs one radix-16 DFT
stage of a Stockham FFT on LDS-resident lines as VALU butterflies (FP32) and
as split-FP16 MFMAs (v_mfma_f32_32x32x16_f16, high and low parts, FP32
accumulation).  Both must equal numpy's DFT after two stages (relative L2
<= 1e-6 for the VALU form, <= 2e-6 for the split form, whose FP16 parts
carry 22 bits, <= 2e-3 for the one-product FP16 form, the speed bound of
any MFMA form); the timed rates of 200 stages are printed and written to
gpurun_out/mfma_stage.json when that folder exists."""

import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2, announce

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "mfma_stage")


def test_mfma_stage_probe(torch_cuda, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/mfma_stage not built"
    announce("MFMA vs VALU radix-16 stage probe")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ld = lambda n: np.fromfile(tmp_path / n, dtype=np.float32)
    x = ld("input.f32").reshape(64, 16, 256, 2)
    x = x[..., 0].astype(np.float64) + 1j * x[..., 1]
    ref = np.fft.fft(np.fft.fft(x, axis=1) / 4, axis=1) / 4
    for name, tol in (("valu.f32", 1e-6), ("mfma.f32", 2e-6),
                      ("mfma16.f32", 2e-3)):
        y = ld(name).reshape(64, 16, 256, 2)
        y = y[..., 0] + 1j * y[..., 1].astype(np.float64)
        assert rel_l2(y, ref) < tol, name
    out = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "mfma_stage.json"), "w") as f:
            json.dump(res, f)
