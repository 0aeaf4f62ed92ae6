"""The C++ facade (include/pickle_gpu.hpp) and pyann.HNSWIndex on the GPU.

The facade is driven by tests/cpp/facade_check as a caller of
pickle::v2::HNSWGraph would drive the reference:
  * Load + AnnSearch on graphs built by the reference: ids and distances
    bit-exact with the reference's own results (golden fixtures);
  * Build with batch 1 and reference rows: the same search results as the
    oracle's single-threaded Build from the same levels;
  * single-query AnnSearch equals the batched AnnSearch.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi
from tests import golden_cases as G
from tests.gpu_util import index_from_oracle
from tests.test_facade import facade_binary

pytestmark = pytest.mark.gpu

DT = {np.dtype(np.uint8): "u8", np.dtype(np.int8): "i8", np.dtype(np.float16): "f16", np.dtype(np.float32): "f32"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def read_out(path, nq, k):
    raw = np.fromfile(path, dtype=np.uint8)
    ids = raw[: nq * k * 4].view(np.int32).reshape(nq, k)
    ds = raw[nq * k * 4:].view(np.float32).reshape(nq, k)
    return ids, ds


@pytest.mark.parametrize("name", ["u8_l2", "i8_ip", "f16_l2", "f32_ip"])
def test_facade_search_matches_reference(tmp_path, name):
    fx, p, base, queries = G.graph_case(name)
    g = G.oracle_graph(fx, p, base)
    say("facade search", name)
    idx = index_from_oracle(g)
    gpath = str(tmp_path / "g.bin")
    idx.save_file(gpath)
    idx.close()
    base.tofile(tmp_path / "base")
    queries.tofile(tmp_path / "q")
    n, nq, dim = base.shape[0], queries.shape[0], base.shape[1]
    for k, ef in fx["searches"][:2]:
        tag = f"k{k}_ef{ef}"
        out = str(tmp_path / f"out_{tag}")
        cmd = [facade_binary(), "search", gpath, str(tmp_path / "base"), str(tmp_path / "q"), str(n), str(nq),
               str(dim), DT[base.dtype], p["metric"], str(p["M"]), str(p["efc"]), str(k), str(ef), out]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        ids, ds = read_out(out, nq, int(k))
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"], err_msg=tag)


@pytest.mark.parametrize("name", ["u8_l2", "f32_ip"])
def test_facade_build_batch1_matches_oracle(tmp_path, name):
    fx, p, base, queries = G.graph_case(name)
    n, nq, dim, M, efc = base.shape[0], queries.shape[0], base.shape[1], p["M"], p["efc"]
    seed = 11
    levels = abi.sample_levels(n, M, seed)
    g = O.Graph(base, G.metric_code(p["metric"]), M, efc, levels)
    g.build(np.arange(n, dtype=np.int32))
    k, ef = 10, 40
    want_ids, want_d, _ = g.search(queries, k, ef)
    base.tofile(tmp_path / "base")
    queries.tofile(tmp_path / "q")
    out, gout = str(tmp_path / "out"), str(tmp_path / "g.bin")
    say("facade build", name, n, dim)
    cmd = [facade_binary(), "build", str(tmp_path / "base"), str(tmp_path / "q"), str(n), str(nq), str(dim),
           DT[base.dtype], p["metric"], str(M), str(efc), str(k), str(ef), str(seed), "1", "1", out, gout]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ids, ds = read_out(out, nq, k)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(ds, want_d)
    assert os.path.exists(gout) and os.path.exists(gout + ".v2meta")


def test_pyann_build_search_save_load(tmp_path):
    from pickle_amd import datasets
    from pickle_amd.pyann import HNSWIndex

    base, queries = datasets.gaussian_mixture(6000, 200, 48, 16, 9, kind="u8")
    gt, _ = O.brute_force(base, queries, 10, O.L2)
    say("pyann build", base.shape)
    idx = HNSWIndex()
    idx.M = 12
    idx.build_ef = 64
    idx.search_ef = 64
    idx.build(base, seed=3)
    assert idx.num_layer >= 2
    ids = idx.search(queries)
    assert ids.shape == (200, 10) and ids.dtype == np.int32
    rec = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)])
    say("  recall@10", rec)
    assert rec > 0.95
    path = str(tmp_path / "pyann.bin")
    idx.save(path)
    idx2 = HNSWIndex()
    idx2.search_ef = 64
    idx2.load(path, data=base)
    np.testing.assert_array_equal(idx2.search(queries), ids)
    with pytest.raises(ValueError):
        idx2.search(queries.astype(np.int8))
    # the fast mode: recall-equivalent to the exact search
    idx2.fast_search = True
    fids = idx2.search(queries)
    frec = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(fids, gt)])
    say("  fast-mode recall@10", frec)
    assert frec >= rec - 0.01
