"""Dataset files the reference's drivers read.

* ``.npy`` (benchmark/bench_util.hpp:85-100, 250-263, via cnpy): base,
  query, label and distance arrays; element types map to pickle's DataType
  as GetNumpyType does (u1 Uint8, i1 Int8, f2 Float16, f4 Float32).
* big-ann binaries (dataset/README.md): ``.fbin`` / ``.u8bin`` / ``.i8bin``
  are ``uint32 n, uint32 d`` then n x d elements; ground-truth files are
  ``uint32 n, uint32 k``, n x k uint32 ids, then n x k float32 distances.

Large files are memory-mapped, so a 100M-row base is not read whole.
"""
from __future__ import annotations

import os

import numpy as np

BIN_DTYPES = {".fbin": np.float32, ".u8bin": np.uint8, ".i8bin": np.int8}
PK_DTYPE = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.float32): 2, np.dtype(np.float16): 3}


def pickle_dtype(dt) -> int:
    """DataType code of a numpy element type (bench_util.hpp:85-100)."""
    dt = np.dtype(dt)
    if dt not in PK_DTYPE:
        raise ValueError(f"unsupported element type {dt} (uint8, int8, float16, float32)")
    return PK_DTYPE[dt]


def _bin_dtype(path, dtype):
    if dtype is not None:
        return np.dtype(dtype)
    ext = os.path.splitext(path)[1]
    if ext not in BIN_DTYPES:
        raise ValueError(f"cannot tell the element type of {path} (.fbin, .u8bin, .i8bin or dtype=)")
    return np.dtype(BIN_DTYPES[ext])


def read_bin(path: str, dtype=None, start: int = 0, count: int | None = None) -> np.ndarray:
    """Rows [start, start + count) of a big-ann vector file (memory-mapped)."""
    dt = _bin_dtype(path, dtype)
    n, d = np.fromfile(path, dtype="<u4", count=2)
    n, d = int(n), int(d)
    size = os.path.getsize(path)
    if size < 8 + n * d * dt.itemsize:
        raise ValueError(f"{path}: header says {n} x {d} but the file holds {size - 8} data bytes")
    count = n - start if count is None else count
    if start < 0 or count < 0 or start + count > n:
        raise ValueError(f"rows [{start}, {start + count}) outside {n}")
    mm = np.memmap(path, dtype=dt, mode="r", offset=8 + start * d * dt.itemsize, shape=(count, d))
    return mm


def write_bin(path: str, rows: np.ndarray) -> None:
    rows = np.ascontiguousarray(rows)
    if rows.ndim != 2:
        raise ValueError("rows must be 2-D")
    with open(path, "wb") as f:
        np.array(rows.shape, dtype="<u4").tofile(f)
        rows.tofile(f)


def read_groundtruth(path: str):
    """(ids uint32 n x k, distances float32 n x k) of a big-ann ground-truth file."""
    n, k = (int(x) for x in np.fromfile(path, dtype="<u4", count=2))
    raw = np.fromfile(path, dtype=np.uint8, offset=8)
    if raw.size != n * k * 8:
        raise ValueError(f"{path}: expected {n * k * 8} bytes after the header, found {raw.size}")
    ids = raw[: n * k * 4].view("<u4").reshape(n, k)
    dists = raw[n * k * 4:].view("<f4").reshape(n, k)
    return ids, dists


def write_groundtruth(path: str, ids: np.ndarray, dists: np.ndarray) -> None:
    ids = np.ascontiguousarray(ids, dtype="<u4")
    dists = np.ascontiguousarray(dists, dtype="<f4")
    if ids.shape != dists.shape or ids.ndim != 2:
        raise ValueError("ids and distances must be n x k")
    with open(path, "wb") as f:
        np.array(ids.shape, dtype="<u4").tofile(f)
        ids.tofile(f)
        dists.tofile(f)


def read_npy(path: str) -> np.ndarray:
    """An .npy array, memory-mapped; nothing in the file is executed."""
    return np.load(path, mmap_mode="r", allow_pickle=False)


def read_vectors(path: str, start: int = 0, count: int | None = None) -> np.ndarray:
    """Rows of a .npy or big-ann vector file, by extension."""
    if path.endswith(".npy"):
        a = read_npy(path)
        if a.ndim != 2:
            raise ValueError(f"{path} is not 2-D")
        pickle_dtype(a.dtype)
        return a[start:None if count is None else start + count]
    return read_bin(path, start=start, count=count)


def recall(ids: np.ndarray, gt_ids: np.ndarray, k: int, dists=None, gt_dists=None) -> float:
    """recall@k.  Without distances, a returned id counts if it is among the
    first k ground-truth labels (benchmark/hnsw_pickle.cpp:94-112).  With
    distances, a returned id whose distance ties the k-th ground-truth
    distance also counts (the big-ann rule, dataset/README.md)."""
    hit = 0
    for i in range(ids.shape[0]):
        truth = set(int(x) for x in gt_ids[i, :k])
        for j in range(k):
            if int(ids[i, j]) in truth:
                hit += 1
            elif dists is not None and gt_dists is not None and dists[i, j] <= gt_dists[i, k - 1]:
                hit += 1
    return hit / (k * ids.shape[0])
