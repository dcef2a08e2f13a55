"""Real-data front end: HDF5 projection I/O and the JSON-configured MBIR run.

The host side reads and writes float32 HDF5 datasets through
``libtomocam_h5.so`` (``include/tomocam_h5.h``).  The device side is
``tc_projections_to_sinogram`` in the HIP library, which turns projections
stored [nproj][nslices][ncols] into sinograms [nslices][nproj][ncols'] and
drops the last column of an even width.  The reference does that transpose
on the host (src/hdf5/reader.h:66-122) and drops the column with
DArray::dropcol (src/dist_array.h:258-282, applied by test/mbir2.cpp:104-107).

``run_json`` follows the configuration schema of the reference's driver
(test/mbir2.cpp:45-74, json/*.json)::

    {"filename": ..., "dataset": ..., "angles": ..., "axis": <centre>,
     "slices": [begin, end],                          (optional)
     "MBIR": {"num_iters": ..., "sigma": ..., "tol": ..., "xtol": ...},
     "output": ...}

It writes the reconstruction to ``output`` as the dataset ``recon``
(test/mbir2.cpp:131-132).
"""

from __future__ import annotations

import ctypes as C
import json
import os
import threading

import numpy as np

LIBNAME = "libtomocam_h5.so"
_HERE = os.path.dirname(os.path.abspath(__file__))
_lock = threading.Lock()
_h5 = None

_c_long_p = C.POINTER(C.c_long)
_c_float_p = C.POINTER(C.c_float)
SIGNATURES = {
    "tc_h5_dims": [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), _c_long_p],
    "tc_h5_read_projections": [C.c_char_p, C.c_char_p, C.c_long, C.c_long,
                               C.c_void_p],
    "tc_h5_read_slices": [C.c_char_p, C.c_char_p, C.c_long, C.c_long,
                          C.c_void_p],
    "tc_h5_read_1d": [C.c_char_p, C.c_char_p, C.c_void_p, C.c_long],
    "tc_h5_write": [C.c_char_p, C.c_char_p, C.c_void_p, C.c_int, _c_long_p,
                    C.c_int],
}


def h5_lib_path() -> str:
    return os.path.join(_HERE, LIBNAME)


def h5_lib():
    """Loads (once) the HDF5 front-end library; raises if it is not built."""
    global _h5
    if _h5 is not None:
        return _h5
    with _lock:
        if _h5 is None:
            path = h5_lib_path()
            if not os.path.exists(path):
                raise RuntimeError(
                    f"{LIBNAME} is not built ({path} missing); run "
                    "make -C tomocam_amd/h5 (needs the HDF5 C library)")
            L = C.CDLL(path)
            L.tc_h5_last_error.restype = C.c_char_p
            L.tc_h5_last_error.argtypes = []
            for name, args in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = C.c_int
                f.argtypes = args
            _h5 = L
    return _h5


def _call(name, *args):
    L = h5_lib()
    st = getattr(L, name)(*args)
    if st != 0:
        msg = L.tc_h5_last_error().decode(errors="replace")
        if st == 1:
            raise ValueError(f"{name}: {msg}")
        raise OSError(f"{name}: {msg}")


def _b(s: str) -> bytes:
    return os.fsencode(s)


def h5_dims(filename: str, dataset: str) -> tuple:
    """Extents of a dataset of rank 1-3."""
    nd = C.c_int()
    dims = (C.c_long * 3)()
    _call("tc_h5_dims", _b(filename), _b(dataset), C.byref(nd), dims)
    return tuple(int(dims[i]) for i in range(nd.value))


def _range(n: int, begin: int, end: int):
    end = n if end is None or end < 0 else end
    if not 0 <= begin <= end <= n:
        raise ValueError(f"slice range [{begin}, {end}) out of bounds "
                         f"({n} slices)")
    return begin, end


def read_projections(filename: str, dataset: str, begin: int = 0,
                     end: int = -1) -> np.ndarray:
    """Slices [begin, end) of a [nproj][nslices][ncols] projection dataset,
    as stored: [nproj][end - begin][ncols] float32."""
    d = h5_dims(filename, dataset)
    if len(d) != 3:
        raise ValueError(f"dataset {dataset} is not 3D")
    begin, end = _range(d[1], begin, end)
    out = np.empty((d[0], end - begin, d[2]), dtype=np.float32)
    _call("tc_h5_read_projections", _b(filename), _b(dataset), begin, end,
          out.ctypes.data_as(C.c_void_p))
    return out


def read_slices(filename: str, dataset: str, begin: int = 0,
                end: int = -1) -> np.ndarray:
    """Slices [begin, end) along axis 0 of a rank-3 dataset."""
    d = h5_dims(filename, dataset)
    if len(d) != 3:
        raise ValueError(f"dataset {dataset} is not 3D")
    begin, end = _range(d[0], begin, end)
    out = np.empty((end - begin, d[1], d[2]), dtype=np.float32)
    _call("tc_h5_read_slices", _b(filename), _b(dataset), begin, end,
          out.ctypes.data_as(C.c_void_p))
    return out


def read_angles(filename: str, dataset: str) -> np.ndarray:
    """A rank-1 float32 dataset (the projection angles)."""
    d = h5_dims(filename, dataset)
    if len(d) != 1:
        raise ValueError(f"dataset {dataset} is not 1D")
    out = np.empty(d[0], dtype=np.float32)
    _call("tc_h5_read_1d", _b(filename), _b(dataset),
          out.ctypes.data_as(C.c_void_p), d[0])
    return out


def write(filename: str, dataset: str, array, truncate: bool = True) -> None:
    """Writes a float32 dataset of rank 1-3; truncate=True creates or
    truncates the file, else the dataset is added."""
    a = np.ascontiguousarray(np.asarray(array, dtype=np.float32))
    if not 1 <= a.ndim <= 3:
        raise ValueError("write: array must have rank 1, 2 or 3")
    dims = (C.c_long * 3)(*a.shape, *([1] * (3 - a.ndim)))
    _call("tc_h5_write", _b(filename), _b(dataset),
          a.ctypes.data_as(C.c_void_p), a.ndim, dims, int(bool(truncate)))


def read_sinogram(filename: str, dataset: str, begin: int = 0,
                  end: int = -1, dropcol: bool = True):
    """Slices [begin, end) of a projection dataset as a device sinogram
    [S, nproj, ncols'] (torch, cuda): the block is read into pinned host
    memory, copied to the GPU and transposed there; with ``dropcol`` an even
    width loses its last column, as the reference's driver does."""
    import torch
    from . import _lib
    proj = read_projections(filename, dataset, begin, end)
    P, S, N = proj.shape
    if S == 0:
        raise ValueError("read_sinogram: empty slice range")
    pinned = torch.from_numpy(proj).pin_memory()
    dproj = pinned.to("cuda", non_blocking=True)
    drop = bool(dropcol and N % 2 == 0)
    sino = torch.empty((S, P, N - int(drop)), dtype=torch.float32,
                       device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.call("tc_projections_to_sinogram", C.c_void_p(dproj.data_ptr()),
              C.c_void_p(sino.data_ptr()), P, S, N, int(drop),
              C.c_void_p(stream))
    torch.cuda.current_stream().synchronize()
    return sino


def run_json(path: str, verbose: bool = False):
    """The reference driver's workflow (test/mbir2.cpp:36-133) on a JSON
    configuration: read, drop a column of even widths, mbir2, write
    ``recon``.  Returns (recon [S, N', N'] numpy, info).  The configuration's
    centre is ``axis`` and is not shifted by the column drop, as in the
    reference; keys it does not read ("p") are ignored as there."""
    from . import gpu
    with open(path) as f:
        cfg = json.load(f)
    for key in ("filename", "dataset", "angles", "axis", "MBIR", "output"):
        if key not in cfg:
            raise ValueError(f"{path}: missing key '{key}'")
    begin, end = 0, -1
    if "slices" in cfg:
        begin, end = int(cfg["slices"][0]), int(cfg["slices"][1])
    params = cfg["MBIR"]
    num_iters = int(params["num_iters"])
    sigma = float(params["sigma"])
    tol = float(params.get("tol", 0.001))
    xtol = float(params.get("xtol", 0.001))
    center = float(int(cfg["axis"]))
    sino = read_sinogram(cfg["filename"], cfg["dataset"], begin, end)
    angles = read_angles(cfg["filename"], cfg["angles"])
    rec, info = gpu.mbir(sino, angles, center, num_iters, sigma, tol, xtol,
                         verbose=verbose)
    out = rec.cpu().numpy()
    write(cfg["output"], "recon", out)
    return out, info


if __name__ == "__main__":
    import sys
    if len(sys.argv) != 2:
        print("usage: python -m tomocam_amd.io <config.json>", file=sys.stderr)
        sys.exit(2)
    _, inf = run_json(sys.argv[1])
    print(inf)
