"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/tomocam_hip.h declares; the ctypes table matches the header."""

import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tomocam_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"^(?:int|const char\*)\s+(tc_\w+)\s*\(([^)]*)\)\s*;",
                         text, flags=re.M):
        args = [a.strip() for a in m.group(2).split(",") if a.strip()]
        if args == ["void"]:
            args = []
        decls[m.group(1)] = args
    return decls


def test_header_parses():
    d = _declared()
    assert "tc_mbir2" in d and "tc_project" in d and len(d) >= 25


def test_library_exports_every_declared_symbol():
    from tomocam_amd import _lib
    lib = _lib.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_ctypes_table_matches_header():
    from tomocam_amd import _lib
    d = _declared()
    for name, args in d.items():
        if name == "tc_last_error":
            continue
        assert name in _lib.SIGNATURES, name
        assert len(_lib.SIGNATURES[name]) == len(args), name


def test_error_path_without_gpu():
    """Argument validation happens before any device call."""
    from tomocam_amd import _lib
    with pytest.raises(ValueError, match="n must be >= 1"):
        _lib.call("tc_set_num_of_gpus", 0)
    _lib.call("tc_set_num_of_gpus", 1)
    with pytest.raises(ValueError, match="sigma"):
        import ctypes as C
        _lib.call("tc_add_tv_hessian", None, 0, -1.0, None)
    with pytest.raises(ValueError, match="stride must be >= 1"):
        _lib.call("tc_ktimer_stride", 0)
    _lib.call("tc_ktimer_stride", 1)


def test_tv_hessian_value_matches_oracle():
    import ctypes as C
    from tomocam_amd import _lib
    from oracle import tomocam_oracle as O
    v = C.c_float()
    _lib.call("tc_add_tv_hessian", None, 0, 0.1, C.byref(v))
    assert abs(v.value - O.tv_hessian_zero(0.1)) <= 1e-6 * v.value


def test_fp64_shape_checks_without_gpu():
    """The double-precision entry points reject bad shapes before touching
    the device (host buffers stand in for device pointers: never used)."""
    import ctypes as C
    from tomocam_amd import _lib
    buf = (C.c_double * 8)()
    p = C.cast(buf, C.c_void_p)
    out = C.c_double()
    with pytest.raises(ValueError, match="FP64 gradient: bad shape"):
        _lib.call("tc_gradient_f64", p, p, p, 0, 16, p, 4, None)
    with pytest.raises(ValueError, match="FP64 function_value: bad shape"):
        _lib.call("tc_function_value_f64", p, p, 1, 1, p, 4, C.byref(out),
                  None)
    with pytest.raises(ValueError, match="xerror: need at least one"):
        _lib.call("tc_xerror_f64", p, p, 0, C.byref(out), None)
    with pytest.raises(ValueError, match="FP64 PSF: bad geometry"):
        h = C.c_void_p()
        _lib.call("tc_psf_create_f64", C.byref(h), p, 0, 16, None)
