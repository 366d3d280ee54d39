"""libdpmring.so loads on the CPU host and exports every symbol include/dpmring.h declares."""
import os
import re

import pytest

from tests.conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dpmring.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dpm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "dpm_create" in syms and "dpm_calc_force_energy" in syms and len(syms) >= 40


def test_library_exports_all_header_symbols():
    from cudaring_amd import dpm
    path = dpm.LIB_PATH
    if not os.path.exists(path):
        pytest.fail("libdpmring.so not built (run __graft_entry__.build())")
    lib = dpm.load_library()
    for s in header_symbols():
        assert hasattr(lib, s), s
    # the Python binding declares exactly the header's entry points
    assert sorted(dpm.SIGNATURES) == header_symbols()
