"""The C++ facade (include/pickle_gpu.hpp) and pyann.HNSWIndex without a GPU:
they compile, link against the C-ABI library, and report argument and
state errors the way the C-ABI does."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "tests", "cpp", "facade_check")


def facade_binary():
    lib = os.path.join(ROOT, "pickle_amd", "libpickle_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libpickle_amd.so not built")
    src = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
    if not os.path.exists(FACADE) or os.path.getmtime(FACADE) < max(
            os.path.getmtime(src), os.path.getmtime(os.path.join(ROOT, "include", "pickle_gpu.hpp"))):
        import __graft_entry__

        __graft_entry__.build_facade_check()
    return FACADE


def test_facade_errors_without_device():
    r = subprocess.run([facade_binary(), "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok 0"


def test_facade_usage_error():
    r = subprocess.run([facade_binary(), "search"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def test_pyann_defaults_and_validation():
    from pickle_amd.pyann import HNSWIndex

    idx = HNSWIndex()
    # pybind/hnsw.cpp:21-25
    assert (idx.search_ef, idx.build_ef, idx.top_k, idx.M, idx.num_layer) == (100, 100, 10, 32, 0)
    idx.search_ef = 40
    idx.top_k = 5
    assert (idx.search_ef, idx.top_k) == (40, 5)
    for name in ("search_ef", "build_ef", "top_k", "M"):
        with pytest.raises(ValueError):
            setattr(idx, name, 0)
    with pytest.raises(RuntimeError):
        idx.search(np.zeros((2, 8), np.uint8))
    with pytest.raises(RuntimeError):
        idx.save("/tmp/never_written")
    with pytest.raises(RuntimeError):
        idx.load("/tmp/never_read")
    with pytest.raises(ValueError):
        idx.build(np.zeros((2, 3, 4), np.uint8))
    with pytest.raises(ValueError):
        idx.build(np.zeros((4, 8), np.int64))
