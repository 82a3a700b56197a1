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


def test_batch_size_limits_are_errors_before_any_device_work():
    lib = _lib.load()
    with open(os.path.join(ROOT, "include", "rainmd.h")) as f:
        limit = int(re.search(r"#define RMD_BATCH_MAX_REPLICAS (\d+)", f.read()).group(1))
    systems = (ctypes.c_void_p * 1)(ctypes.c_void_p(1))  # never read: the count is refused first
    out = ctypes.c_void_p()
    for count, message in ((0, b"at least one replica"), (limit + 1, b"at most %d replicas" % limit)):
        rc = lib.rmd_batch_create(ctypes.cast(systems, ctypes.c_void_p), count, None, None, ctypes.byref(out))
        assert rc != 0
        assert message in lib.rmd_last_global_error()
        assert not out.value
