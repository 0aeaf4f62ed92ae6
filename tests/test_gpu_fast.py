"""Fast-mode search (pk_search_fast_device) against the exact search and
the oracle.

The fast mode keeps the reference's accept and stopping rules
(include/hnsw.hpp:398-434) but orders ties by id and may expand several
entries per round, so its bar is:
  * width 1 on float graphs (no ties) returns the exact search's ids and
    distances, which are the reference's (bit-exact, golden fixtures);
  * results do not depend on the LDS visited filter's size (a forgotten id
    is only evaluated again), only the distance count does;
  * with ef >= N every node is reachable and the result is the brute-force
    top-k;
  * recall@k of every width stays within 0.01 of the exact search's on the
    golden graphs, and uneven per-query k / ef are honoured.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi
from tests import golden_cases as G
from tests.gpu_util import index_from_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def graphs():
    cache = {}

    def get(name):
        if name not in cache:
            fx, p, base, queries = G.graph_case(name)
            g = G.oracle_graph(fx, p, base)
            cache[name] = (fx, p, base, queries, g)
        return cache[name]

    return get


def recall(ids, gt, k):
    return float(np.mean([len(set(ids[i, :k]) & set(gt[i, :k])) / k for i in range(len(ids))]))


# unrounded float data (f32_l2's vectors are integer-valued: its distances tie)
@pytest.mark.parametrize("name", ["f32_ip", "f16_l2", "f16_ip"])
def test_width1_equals_exact_on_float_graphs(graphs, name):
    fx, p, base, queries, g = graphs(name)
    idx = index_from_oracle(g)
    for k, ef in fx["searches"]:
        k, ef = int(k), int(ef)
        say("case", name, "k", k, "ef", ef)
        want_ids, want_d, want_n, want_c = idx.search(queries, k, ef, counters=True)
        idx.set_fast_tuning(width=1, hash_bits=14)
        ids, ds, n, c = idx.search_fast(queries, k, ef, counters=True)
        np.testing.assert_array_equal(n, want_n)
        np.testing.assert_array_equal(ids, want_ids)
        np.testing.assert_array_equal(ds, want_d)
        # the same expansions (hops) in every layer
        np.testing.assert_array_equal(c["hops"], want_c["hops"])
    idx.close()


@pytest.mark.parametrize("name", ["f32_ip", "f16_l2", "u8_l2", "i8_ip"])
def test_results_independent_of_filter_size(graphs, name):
    fx, p, base, queries, g = graphs(name)
    idx = index_from_oracle(g)
    for width in (1, 2, 4):
        idx.set_fast_tuning(width=width, hash_bits=14)
        big = idx.search_fast(queries, 10, 64, counters=True)
        idx.set_fast_tuning(width=width, hash_bits=8)
        small = idx.search_fast(queries, 10, 64, counters=True)
        say("case", name, "width", width, "dists", int(big[3]["dists"].sum()), int(small[3]["dists"].sum()))
        if name in ("f32_ip", "f16_l2"):
            # no ties: identical results and expansions
            np.testing.assert_array_equal(small[0], big[0])
            np.testing.assert_array_equal(small[1], big[1])
            np.testing.assert_array_equal(small[3]["hops"], big[3]["hops"])
        else:
            # integer distances tie; the merge grouping can change which of
            # two equal-distance ids stays, never a returned distance
            np.testing.assert_array_equal(small[1], big[1])
        assert (small[3]["dists"][:, 0] >= big[3]["dists"][:, 0]).all()
        # never the same id twice
        for row, cnt in zip(small[0], small[2]):
            assert len(set(row[:cnt])) == cnt
    idx.close()


@pytest.mark.parametrize("name", ["f32_ip", "f16_l2"])
def test_results_independent_of_filter_tags(graphs, name):
    """The filter keeps b - slot_bits bits of each hashed id per entry, b
    the bits of the index capacity: none for a 3,000-row index, up to 15
    for a capacity of 2^24 + 1 (25-bit ids).  The same graph searched
    through both returns the same results at every filter size (tie-free
    data)."""
    fx, p, base, queries, g = graphs(name)
    small = index_from_oracle(g)
    big = index_from_oracle(g, capacity=(1 << 24) + 1)
    for hb in (8, 11, 14):
        say("case", name, "hash_bits", hb)
        small.set_fast_tuning(width=2, hash_bits=hb)
        big.set_fast_tuning(width=2, hash_bits=hb)
        a = small.search_fast(queries, 10, 64)
        b = big.search_fast(queries, 10, 64)
        np.testing.assert_array_equal(b[2], a[2])
        np.testing.assert_array_equal(b[0], a[0])
        np.testing.assert_array_equal(b[1], a[1])
    small.close()
    big.close()


@pytest.mark.parametrize("name", G.GRAPH_CASES)
def test_ef_covering_graph_is_brute_force(name):
    """A 400-node oracle graph of each golden case's first rows, searched
    at ef 512 >= N: every node is reached, so the result is the brute force."""
    fx, p, base, queries = G.graph_case(name)
    n = 400
    base = np.ascontiguousarray(base[:n])
    metric = G.metric_code(p["metric"])
    g = O.Graph(base, metric, p["M"], p["efc"], O.sample_levels(n, p["M"], seed=9))
    g.build(np.arange(n, dtype=np.int32))
    idx = index_from_oracle(g)
    gt_ids, gt_d = O.brute_force(base, queries, 10, metric)
    ef = 512
    for width in (1, 4):
        idx.set_fast_tuning(width=width, hash_bits=10)
        ids, ds, cnt = idx.search_fast(queries, 10, ef)
        assert (cnt == 10).all()
        # distances equal the brute force's (FP64 sums, rounded to FP32 by
        # the comparison's tolerance for floats; exact for integers)
        if base.dtype.kind in "ui":
            np.testing.assert_array_equal(ds, gt_d.astype(np.float32))
        else:
            np.testing.assert_allclose(ds, gt_d, rtol=1e-5, atol=1e-4)
    idx.close()


@pytest.mark.parametrize("name", G.GRAPH_CASES)
def test_recall_matches_exact(graphs, name):
    fx, p, base, queries, g = graphs(name)
    idx = index_from_oracle(g)
    metric = G.metric_code(p["metric"])
    gt_ids, _ = O.brute_force(base, queries, 10, metric)
    for ef in (10, 20, 40):
        ex_ids, _, _ = idx.search(queries, 10, ef)
        r_ex = recall(ex_ids, gt_ids, 10)
        for width in (1, 2, 4):
            idx.set_fast_tuning(width=width)
            ids, ds, cnt = idx.search_fast(queries, 10, ef)
            r = recall(ids, gt_ids, 10)
            say("case", name, "ef", ef, "width", width, "recall", round(r, 4), "exact", round(r_ex, 4))
            assert r >= r_ex - 0.01, (name, ef, width, r, r_ex)
            # ascending distances
            assert (np.diff(ds, axis=1) >= 0).all()
    idx.close()


def test_uneven_k_and_ef(graphs):
    fx, p, base, queries, g = graphs("u8_l2")
    idx = index_from_oracle(g)
    rng = np.random.default_rng(5)
    nq = len(queries)
    ks = rng.integers(1, 21, nq).astype(np.int32)
    efs = np.maximum(rng.integers(8, 300, nq), ks).astype(np.int32)
    idx.set_fast_tuning(width=2)
    ids, ds, cnt = idx.search_fast(queries, ks, efs, k_stride=20)
    for i in range(nq):
        one_ids, one_d, one_n = idx.search_fast(queries[i:i + 1], int(ks[i]), int(efs[i]), k_stride=20)
        assert cnt[i] == one_n[0] == ks[i]
        np.testing.assert_array_equal(ids[i], one_ids[0])
        np.testing.assert_array_equal(ds[i], one_d[0])
        assert (ids[i, ks[i]:] == -1).all()
    idx.close()


def test_fast_argument_errors(graphs):
    fx, p, base, queries, g = graphs("u8_l2")
    idx = index_from_oracle(g)
    with pytest.raises(abi.PickleError):
        idx.search_fast(queries, 10, 600)  # ef above 512
    with pytest.raises(abi.PickleError):
        idx.set_fast_tuning(width=5)
    with pytest.raises(abi.PickleError):
        idx.set_fast_tuning(hash_bits=20)
    idx.close()
