"""The C-ABI library loads and exports every symbol include/nexus_mpm.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from nexus_amd import mpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nexus_mpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nxm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_wrapper_agree():
    assert declared_symbols() == sorted(mpm.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = mpm.library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_version_string():
    assert b"gfx950" in mpm.library().nxm_version()
