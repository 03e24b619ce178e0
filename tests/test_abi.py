"""The C-ABI library loads without a GPU and exports exactly what include/akx.h
declares; the Python binding covers every entry point.  No compute call."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "akx.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(akx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "akx_solve_step" in names and "akx_init" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from akantu_amd import akx
    lib = ctypes.CDLL(akx.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_covers_every_declared_symbol():
    from akantu_amd import akx
    assert sorted(akx.SIGNATURES) == declared()


def test_no_device_raises_cleanly():
    from akantu_amd import akx
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(akx.AkxError):
        akx.Context(0)
