"""C-ABI library: loads and exports every declared symbol (no GPU needed)."""
import os
import re
import subprocess

from pickle_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pickle_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pk_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (pk_\w+)", out))
    for name in declared_symbols():
        assert name in exported, name
        assert getattr(L, name) is not None
    assert L.pk_version() >= 100


def test_errors_without_device_are_reported_not_raised_as_crash():
    # argument checks run before any HIP call
    import ctypes as C
    h = C.c_void_p()
    rc = abi.lib().pk_index_create(7, 1, 16, 8, 10, 100, C.byref(h))
    assert rc == -1 and b"dtype" in abi.lib().pk_last_error()
    rc = abi.lib().pk_index_create(0, 5, 16, 8, 10, 100, C.byref(h))
    assert rc == -1 and b"metric" in abi.lib().pk_last_error()
