"""The committed fixtures' inputs are reproducible on this host: the
1024^2 golden's sinogram (oracle projector + seeded noise) matches the
checksum stored with it."""

import os
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_mbir2_golden as G  # noqa: E402


def test_mbir2_1024_golden_input_reproducible():
    _, sino = G.make_sinogram(1024)
    gold = np.load(G.fixture_path(1024))
    np.testing.assert_allclose([sino.sum(), np.linalg.norm(sino)],
                               gold["sino_checksum"], rtol=1e-12)
    assert gold["sampled"].shape == (64, 64)
