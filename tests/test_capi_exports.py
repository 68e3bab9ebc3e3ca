"""The C-ABI library loads on a machine without a GPU and exports every entry
point declared in include/plis.h (no compute calls here)."""
import ctypes as C
import os
import re

from pliscuda_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "plis.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plis_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert set(declared()) == set(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    missing = [f for f in declared() if not hasattr(lib, f)]
    assert not missing, missing


def test_create_fails_cleanly_without_device_or_with_bad_model():
    """plis_create returns an error code (never exits) for an invalid model."""
    L = capi.lib()
    net = capi.Network(0, 0, 1, None, None, None, None)
    h = C.c_void_p()
    rc = L.plis_create(C.byref(net), 1, None, None, None, None, C.byref(h))
    assert rc < 0 and not h.value
    assert L.plis_last_error(None)
