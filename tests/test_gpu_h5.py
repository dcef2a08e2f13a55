"""Real-data front end on the GPU: the projections -> sinogram transpose
with the even-width column drop (bit-exact against numpy), and the JSON
driver end to end against a direct mbir2 call on the same data."""

import json
import os

import numpy as np
import pytest

from tomocam_amd import phantom as PH
from conftest import announce

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(
                  os.environ.get("HDF5_DIR", "/opt/conda") + "/include/hdf5.h"),
                  reason="no HDF5 C library on this machine")]


@pytest.mark.parametrize("P,S,N,drop", [(7, 3, 6, 1), (7, 3, 6, 0),
                                        (5, 1, 9, 0), (180, 4, 256, 1)])
def test_projections_to_sinogram(torch_cuda, P, S, N, drop):
    torch = torch_cuda
    import ctypes as C
    from tomocam_amd import _lib
    proj = np.random.default_rng(P + S + N).standard_normal(
        (P, S, N)).astype(np.float32)
    d = torch.from_numpy(proj).cuda()
    out = torch.empty((S, P, N - drop), device="cuda")
    announce(f"projections_to_sinogram P={P} S={S} N={N} drop={drop}")
    _lib.call("tc_projections_to_sinogram", C.c_void_p(d.data_ptr()),
              C.c_void_p(out.data_ptr()), P, S, N, drop, None)
    torch.cuda.synchronize()
    ref = np.ascontiguousarray(proj.transpose(1, 0, 2)[:, :, :N - drop])
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_run_json_matches_direct_mbir(torch_cuda, tmp_path):
    """A 3-slice projection file of width 64 (even: the driver drops the last
    column), slices [1, 3) selected, centre 'axis' 32, 5 iterations."""
    torch = torch_cuda
    from tomocam_amd import gpu, io
    N, P, S = 64, 48, 3
    ang = PH.angles(P)
    img = torch.tensor(np.stack([PH.random_ellipses(N, 90 + s)
                                 for s in range(S)]), dtype=torch.float32).cuda()
    announce("run_json end to end")
    sino = gpu.project(img, ang)                       # [S, P, N]
    proj = sino.permute(1, 0, 2).contiguous().cpu().numpy()  # as stored
    f = str(tmp_path / "scan.h5")
    io.write(f, "projs", proj)
    io.write(f, "angs", ang.astype(np.float32), truncate=False)
    out_file = str(tmp_path / "rec.h5")
    cfg = {"filename": f, "dataset": "projs", "angles": "angs", "axis": 32,
           "slices": [1, 3],
           "MBIR": {"num_iters": 5, "tol": 1e-6, "xtol": 1e-6, "sigma": 100.0,
                    "p": 1.2},
           "output": out_file}
    cpath = str(tmp_path / "cfg.json")
    with open(cpath, "w") as fh:
        json.dump(cfg, fh)
    rec, info = io.run_json(cpath)
    assert rec.shape == (2, N - 1, N - 1)
    direct, _ = gpu.mbir(sino[1:3, :, :N - 1].contiguous(), ang, 32.0, 5,
                         100.0)
    np.testing.assert_array_equal(rec, direct.cpu().numpy())
    np.testing.assert_array_equal(io.read_slices(out_file, "recon"), rec)
    assert info["iterations"] == 5
