"""The committed fixtures' inputs are reproducible on this host: the
1024^2 golden's sinogram (oracle projector + seeded noise) matches the
checksum stored with it."""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_mbir2_golden as G  # noqa: E402


def test_mbir2_1024_golden_input_reproducible():
    _, sino = G.make_sinogram(1024)
    gold = np.load(G.fixture_path(1024))
    np.testing.assert_allclose([sino.sum(), np.linalg.norm(sino)],
                               gold["sino_checksum"], rtol=1e-12)
    assert gold["sampled"].shape == (64, 64)


@pytest.mark.parametrize("N", [1024, 2048])
def test_mbir2_100it_fixture_fields(N):
    """The 100-iteration fixtures (the workloads' length; the 2048^2 one is
    the benchmark's FP64 accuracy check) come from the same inputs as the
    30-iteration ones and carry the fields the benchmark reads."""
    a = np.load(G.fixture_path(N))
    b = np.load(G.fixture_path(N, 100))
    assert int(b["iterations"]) == 100 and b["sampled"].shape == (64, 64)
    np.testing.assert_allclose(a["sino_checksum"], b["sino_checksum"],
                               rtol=1e-12)
    assert abs(float(a["lipschitz"]) - float(b["lipschitz"])) <= 1e-12
    assert float(b["step"]) == 1.0
