"""The C-ABI library loads and exports every entry point include/*.h
declares; argument validation happens before any device call."""
import ctypes as C
import glob
import os
import re

import pytest

import zola_hip as Z

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zola_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(Z.LIB_PATH):
        pytest.fail("libzola_hip.so not built (run zola-3d_amd/build.py)")
    lib = C.CDLL(Z.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(Z.SYMBOLS) <= set(names)


def test_invalid_arguments_are_rejected_before_device_use():
    lib = Z.lib()
    h = C.c_void_p()
    bad = Z.optics(size_fft=0)
    assert lib.zola_ctx_create(C.byref(bad), 21, None, 0, C.byref(h)) == -1
    assert b"invalid optics" in lib.zola_last_error(h)
    lib.zola_ctx_destroy(h)
    h = C.c_void_p()
    big = Z.optics(size_fft=512)
    assert lib.zola_ctx_create(C.byref(big), 21, None, 0, C.byref(h)) == -2
    assert b"pupil grid" in lib.zola_last_error(h)
    lib.zola_ctx_destroy(h)
    h = C.c_void_p()
    assert lib.zola_ctx_create(C.byref(Z.optics()), 23, None, 0, C.byref(h)) == -2
    assert b"ROI side" in lib.zola_last_error(h)
    lib.zola_ctx_destroy(h)
    with pytest.raises(Z.ZolaError):
        Z.Context(Z.optics(size_fft=64, xystep=-1), 21, [0.0])
