"""The C ABI on the CPU: the library loads, exports every function that
include/*.h declares, and reports errors without a GPU (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

import polyspectrum_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("ps_ctx_create", "ps_polyspectrum", "ps_batch_create", "polyspectrum", "num_polygons"):
        assert n in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_loads():
    L = P.load_library()
    assert L.ps_strerror(0) == b"ok"
    assert b"unsupported" in L.ps_strerror(-2)
    for n in (16, 24, 32, 48, 64, 96, 128, 144, 192, 216, 256, 384, 512):
        assert L.ps_grid_supported(n) == 1
    for n in (0, 10, 20, 1024, 640):
        assert L.ps_grid_supported(n) == 0


def test_errors_without_device_or_bad_arguments():
    L = P.load_library()
    h = ctypes.c_void_p()
    assert L.ps_ctx_create(20, 100.0, 1, 1, ctypes.byref(h)) == -2     # unsupported N
    assert L.ps_ctx_create(32, -1.0, 1, 1, ctypes.byref(h)) == -1      # bad box size
    assert L.ps_ctx_create(32, 100.0, 1, 0, ctypes.byref(h)) == -1     # no slot
    if P.device_count() == 0:
        assert L.ps_ctx_create(32, 100.0, 1, 1, ctypes.byref(h)) == -6  # no GPU here
        with pytest.raises(P.GPUError):
            P.Context(32, 100.0)
    assert L.ps_ctx_destroy(None) == -1
    assert L.ps_batch_destroy(None) == -1
