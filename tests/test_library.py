"""The C-ABI library builds for gfx950, loads, and exports every entry point include/uvh.h
declares (no device needed; nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

from uasource_amd import uvh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(REPO, "include", "uvh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uvhb?_[a-z_]+)\s*\(", text)))


def test_header_and_binding_list_agree():
    from uasource_amd.batch import EXPORTED_BATCH
    from uasource_amd.linsolve import EXPORTED_LIN
    assert declared_entry_points() == sorted(uvh.EXPORTED + EXPORTED_BATCH + EXPORTED_LIN)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(uvh.LIB_PATH):
        pytest.skip("libuvh.so not built (python -m uasource_amd.csrc.build)")
    lib = ctypes.CDLL(uvh.LIB_PATH)
    for name in declared_entry_points():
        assert hasattr(lib, name), name


def test_library_contains_gfx950_code_object():
    if not os.path.exists(uvh.LIB_PATH):
        pytest.skip("libuvh.so not built")
    blob = open(uvh.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_unsupported_options_are_refused_before_any_device_work():
    from uasource_amd.ctrlvar import default_ctrlvar
    cv = default_ctrlvar()
    uvh.check_supported(cv)
    for law in ("Budd", "W-N0", "Coulomb", "C", "Tsai", "minCW-N0", "Cornford", "rpCW-N0", "Umbi", "rCW-N0",
                "Joughin", "rCW-v0"):
        c = cv.copy()
        c.SlidingLaw = law
        uvh.check_supported(c)
    for attr, val in [("SlidingLaw", "Schoof"), ("FlowApproximation", "SSHEET"),
                      ("ThicknessPenalty", True)]:
        c = cv.copy()
        setattr(c, attr, val)
        with pytest.raises(uvh.UvhUnsupported):
            uvh.check_supported(c)
    c = cv.copy()
    c.Development.Pre2025uvhAssembly = True
    with pytest.raises(uvh.UvhUnsupported):
        uvh.check_supported(c)
