"""The C ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fvp_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"fvp_create", "fvp_destroy", "fvp_residual", "fvp_begin_step", "fvp_jv"} <= names


def test_library_exports_every_declared_symbol():
    from fvmpor_amd import lib as fl
    L = ctypes.CDLL(fl.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from fvmpor_amd import lib as fl
    assert _declared() <= set(fl.SIGNATURES)


def test_last_error_without_device():
    from fvmpor_amd import lib as fl
    L = fl.lib()
    assert isinstance(L.fvp_last_error(), bytes)
    # invalid arguments are rejected before any device call
    rc = L.fvp_create(None, 0, ctypes.byref(ctypes.c_void_p()))
    assert rc < 0 and b"null" in L.fvp_last_error()
