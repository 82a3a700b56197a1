"""The C-ABI library loads and exports every entry point include/rainmd.h
declares (no compute without a GPU)."""
import ctypes
import os
import re

import rainmd_amd
from rainmd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "rainmd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rmd_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIBRARY_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_abi_version():
    assert _lib.load().rmd_abi_version() == 1


def test_device_query_without_a_gpu_is_an_error_not_a_crash():
    info = rainmd_amd.device_info()
    if info["device_count"] == 0:
        assert not info["ok"]
