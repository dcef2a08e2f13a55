"""HDF5 front end, host side (libtomocam_h5.so; no GPU): the library
exports what include/tomocam_h5.h declares, and float32 datasets round-trip
through write / dims / read with the reference's layouts
(src/hdf5/reader.h, src/hdf5/writer.h)."""

import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tomocam_h5.h")
HAVE_HDF5 = os.path.exists(os.environ.get("HDF5_DIR", "/opt/conda")
                           + "/include/hdf5.h")
pytestmark = pytest.mark.skipif(not HAVE_HDF5,
                                reason="no HDF5 C library on this machine")


def test_library_exports_declared_symbols():
    from tomocam_amd import io
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = re.findall(r"^(?:int|const char\*)\s+(tc_h5_\w+)\s*\(", text,
                       flags=re.M)
    assert len(names) == 6
    lib = io.h5_lib()
    assert [n for n in names if not hasattr(lib, n)] == []
    assert set(io.SIGNATURES) == set(names) - {"tc_h5_last_error"}


def test_round_trip_and_hyperslabs(tmp_path):
    from tomocam_amd import io
    f = str(tmp_path / "data.h5")
    rng = np.random.default_rng(1)
    proj = rng.standard_normal((7, 5, 6)).astype(np.float32)  # [P][S][N]
    ang = np.linspace(0, np.pi, 7, endpoint=False).astype(np.float32)
    io.write(f, "projs", proj)
    io.write(f, "extra", proj[:2], truncate=False)
    io.write(f, "angs", ang, truncate=False)
    assert io.h5_dims(f, "projs") == (7, 5, 6)
    assert io.h5_dims(f, "angs") == (7,)
    np.testing.assert_array_equal(io.read_angles(f, "angs"), ang)
    np.testing.assert_array_equal(io.read_projections(f, "projs"), proj)
    np.testing.assert_array_equal(io.read_projections(f, "projs", 1, 4),
                                  proj[:, 1:4])
    np.testing.assert_array_equal(io.read_slices(f, "projs", 2, 5), proj[2:5])
    np.testing.assert_array_equal(io.read_slices(f, "extra"), proj[:2])
    assert io.read_projections(f, "projs", 3, 3).shape == (7, 0, 6)
    # truncate=True starts a new file
    io.write(f, "projs", proj[:1])
    with pytest.raises(OSError, match="cannot open dataset"):
        io.h5_dims(f, "angs")


def test_errors(tmp_path):
    from tomocam_amd import io
    f = str(tmp_path / "e.h5")
    io.write(f, "a", np.zeros((2, 3, 4), np.float32))
    with pytest.raises(OSError, match="cannot open HDF5 file"):
        io.h5_dims(str(tmp_path / "missing.h5"), "a")
    with pytest.raises(OSError, match="cannot open dataset"):
        io.h5_dims(f, "nope")
    with pytest.raises(ValueError, match="out of bounds"):
        io.read_projections(f, "a", 1, 4)
    with pytest.raises(ValueError, match="not 1D"):
        io.read_angles(f, "a")
    with pytest.raises(ValueError, match="rank 1, 2 or 3"):
        io.write(f, "b", np.zeros((1, 1, 1, 1), np.float32))
    with pytest.raises(ValueError, match="not 3D"):
        io.write(f, "c", np.zeros(3, np.float32), truncate=False)
        io.read_projections(f, "c")
