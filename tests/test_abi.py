"""The C-ABI library loads and exports every entry point include/pvglu.h
declares (no compute: runs without a GPU)."""
import os
import re

import phdthesis_albertojimenezruiz_amd as pv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pvglu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pvglu_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = pv.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_device_count_without_gpu():
    L = pv.lib()
    assert b"gfx950" in L.pvglu_version()
    assert L.pvglu_device_count() >= 0


def test_library_has_gfx950_code_object():
    data = open(pv.LIB_PATH, "rb").read()
    assert b"gfx950" in data
