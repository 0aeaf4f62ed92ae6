"""Dataset files: big-ann binaries, ground truth, .npy (CPU)."""
import os

import numpy as np
import pytest

from pickle_amd import formats


def test_bin_layout_matches_the_published_format(tmp_path):
    rows = np.arange(12, dtype=np.uint8).reshape(3, 4)
    p = str(tmp_path / "x.u8bin")
    formats.write_bin(p, rows)
    raw = open(p, "rb").read()
    # uint32 n, uint32 d, little-endian, then the rows one after another
    assert raw[:8] == bytes([3, 0, 0, 0, 4, 0, 0, 0])
    assert raw[8:] == bytes(range(12))
    np.testing.assert_array_equal(formats.read_bin(p), rows)
    np.testing.assert_array_equal(formats.read_bin(p, start=1, count=2), rows[1:])


@pytest.mark.parametrize("ext,dt", [(".fbin", np.float32), (".u8bin", np.uint8), (".i8bin", np.int8)])
def test_bin_round_trip(tmp_path, ext, dt):
    rng = np.random.default_rng(1)
    rows = (rng.normal(size=(100, 24)) * 40).astype(dt)
    p = str(tmp_path / f"v{ext}")
    formats.write_bin(p, rows)
    got = formats.read_vectors(p)
    assert got.dtype == dt
    np.testing.assert_array_equal(got, rows)


def test_bin_errors(tmp_path):
    p = str(tmp_path / "short.fbin")
    with open(p, "wb") as f:
        np.array([10, 4], "<u4").tofile(f)
        np.zeros(8, np.float32).tofile(f)
    with pytest.raises(ValueError):
        formats.read_bin(p)
    with pytest.raises(ValueError):
        formats.read_bin(str(tmp_path / "v.bin"))  # unknown extension


def test_groundtruth_round_trip(tmp_path):
    ids = np.array([[3, 1, 2], [0, 5, 4]], np.uint32)
    ds = np.array([[0.5, 1.0, 2.0], [0.0, 3.0, 3.0]], np.float32)
    p = str(tmp_path / "gt.ibin")
    formats.write_groundtruth(p, ids, ds)
    assert os.path.getsize(p) == 8 + 2 * 3 * 8
    gi, gd = formats.read_groundtruth(p)
    np.testing.assert_array_equal(gi, ids)
    np.testing.assert_array_equal(gd, ds)


def test_npy_and_dtype_codes(tmp_path):
    p = str(tmp_path / "b.npy")
    a = np.arange(20, dtype=np.float16).reshape(5, 4)
    np.save(p, a)
    np.testing.assert_array_equal(formats.read_vectors(p, 1, 2), a[1:3])
    # bench_util.hpp:85-100 (DataType values of include/common.hpp:21-26)
    assert [formats.pickle_dtype(t) for t in (np.uint8, np.int8, np.float32, np.float16)] == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        formats.pickle_dtype(np.int64)


def test_recall_label_and_tie_rules():
    gt = np.array([[1, 2, 3]])
    gd = np.array([[1.0, 2.0, 3.0]], np.float32)
    ids = np.array([[1, 2, 9]])
    ds = np.array([[1.0, 2.0, 3.0]], np.float32)  # 9 ties the 3rd distance
    assert formats.recall(ids, gt, 3) == pytest.approx(2 / 3)
    assert formats.recall(ids, gt, 3, ds, gd) == pytest.approx(1.0)
