"""libdoug_hip.so loads and exports every entry point include/doug.h declares
(no compute call is made here)."""
import ctypes
import os
import re

from conftest import ROOT
from doug_amd import _lib


def _declared():
    text = open(os.path.join(ROOT, "include", "doug.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(doug_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "doug_pcg" in names and "doug_setup" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) == set(_lib.SIGNATURES), set(_declared()) ^ set(_lib.SIGNATURES)


def test_fortran_module_binds_every_entry_point():
    # doug_capi_mod.f90 is the ISO_C_BINDING module a Fortran caller (DOUG's
    # driver) uses: one bind(C) interface per header entry point
    text = open(os.path.join(ROOT, "doug_amd", "fortran", "doug_capi_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(doug_[a-z_0-9]+)"\)', text))
    assert bound == set(_declared()), bound ^ set(_declared())
