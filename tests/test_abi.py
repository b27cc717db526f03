"""The C-ABI library loads and exports every entry point the public header
declares (no compute: these run without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aperture_amd.h")
LIB = os.path.join(ROOT, "aperture3_amd", "libaperture_amd.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(apt_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "apt_create" in names and "apt_ptc_update" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libaperture_amd.so is not built (run aperture3_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (apt_[a-z0-9_]+)$", out, re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    import aperture3_amd as apt
    lib = apt.load_library()
    for name in declared():
        assert hasattr(lib, name)
    assert set(apt.SIGNATURES) == set(declared())


def test_create_rejects_bad_grid_without_gpu_work():
    # argument validation happens before any device work
    import ctypes as C
    import aperture3_amd as apt
    lib = apt.load_library()
    h = C.c_void_p()
    assert lib.apt_create(None, None, C.byref(h)) == 1
    assert lib.apt_destroy(None) == 1
