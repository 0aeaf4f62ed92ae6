"""Vectors read from dataset files (.npy, big-ann .fbin / .u8bin / .i8bin,
benchmark/bench_util.hpp:102-130, dataset/README.md:8-10) through the GPU
search: the files' rows build a graph (the oracle), the GPU searches it with
the file's queries through pk_search, and the results equal the oracle's
search of the same arrays; a written ground-truth file reads back and gives
the recall the reference's definition gives (hnsw_pickle.cpp:94-112)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import datasets, formats
from tests.gpu_util import index_from_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("ext", [".npy", ".fbin", ".u8bin", ".i8bin"])
def test_file_vectors_through_the_gpu_search(tmp_path, ext):
    n, nq, dim, M, efc = 4000, 100, 96, 16, 64
    b, q = datasets.gaussian_mixture(n, nq, dim, 16, 51, kind="u8")
    if ext == ".fbin":
        b, q = b.astype(np.float32), q.astype(np.float32)
    elif ext == ".i8bin":
        b, q = (b.astype(np.int16) - 128).astype(np.int8), (q.astype(np.int16) - 128).astype(np.int8)
    bp, qp = str(tmp_path / f"base{ext}"), str(tmp_path / f"query{ext}")
    if ext == ".npy":
        np.save(bp, b)
        np.save(qp, q)
    else:
        formats.write_bin(bp, b)
        formats.write_bin(qp, q)
    base = np.ascontiguousarray(formats.read_vectors(bp))
    queries = np.ascontiguousarray(formats.read_vectors(qp))
    assert base.dtype == b.dtype and np.array_equal(base, b) and np.array_equal(queries, q)
    say("file", ext, base.dtype, base.shape)
    g = O.Graph(base, O.L2, M, efc, O.sample_levels(n, M, seed=51))
    g.build(np.arange(n, dtype=np.int32))
    idx = index_from_oracle(g)
    ids, ds, cnt = idx.search(queries, 10, 48)
    w_ids, w_d, w_n = g.search(queries, 10, 48)
    np.testing.assert_array_equal(ids, w_ids)
    np.testing.assert_array_equal(ds, w_d)
    # ground truth written in the big-ann layout and read back
    gt_ids, gt_d = O.brute_force(base, queries, 10, O.L2)
    gp = str(tmp_path / "gt.ibin")
    formats.write_groundtruth(gp, gt_ids, gt_d.astype(np.float32))
    r_ids, r_d = formats.read_groundtruth(gp)
    np.testing.assert_array_equal(r_ids, gt_ids)
    r = formats.recall(ids, r_ids, 10)
    want = np.mean([len(set(a) & set(bb)) / 10 for a, bb in zip(ids, gt_ids)])
    assert abs(r - want) < 1e-12 and r > 0.8
    idx.close()
