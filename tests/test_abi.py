"""The C ABI: libbia_hip.so loads (no GPU needed) and exports every entry point declared in
include/*.h; the Python binding declares exactly that set; the Fortran host side built."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:void|int|char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_reference_symbols():
    names = declared_functions()
    for s in ("find_dxyuv_kernel_", "invariance_gpu_", "rkf_c_", "sincos_"):
        assert s in names


def test_library_exports_every_declared_symbol():
    from bia_3d_gpu_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header_exactly():
    from bia_3d_gpu_amd import _lib
    assert set(_lib.declared_symbols()) == declared_functions()
    _lib.load()  # sets every restype/argtypes; raises on a missing export


def test_sincos_symbol_matches_libm():
    import math
    from bia_3d_gpu_amd import _lib
    L = _lib.load()
    for x in (0.0, 0.3, -2.5, 100.0):
        xv, s, c = ctypes.c_double(x), ctypes.c_double(), ctypes.c_double()
        L.sincos_(ctypes.byref(xv), ctypes.byref(s), ctypes.byref(c))
        assert s.value == math.sin(x) and c.value == math.cos(x)


def test_error_strings_without_gpu():
    from bia_3d_gpu_amd import _lib
    L = _lib.load()
    assert L.bia_error_string(0) == b"ok"
    assert L.bia_error_string(1) == b"invalid argument"
    assert L.bia_version() == 10000


def test_fortran_driver_built():
    assert os.path.exists(os.path.join(ROOT, "bia_3d_gpu_amd", "fortran", "bia_run"))


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference checkout absent")
def test_reference_host_programs_link_unmodified():
    for b in ("bia_gpu_alg", "dim_gpu"):
        assert os.path.exists(os.path.join(ROOT, "oracle", "_ref", b))
