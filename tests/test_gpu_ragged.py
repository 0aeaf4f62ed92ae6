"""Search at ragged dimensions and on tiny graphs, through the C-ABI.

The golden graphs have dims that are multiples of 16, so they never run the
float distance's dim % 16 tail (include/distance/avx512.hpp's remainder loop
as the reference build compiles it: an 8-wide block, a 4-wide block, scalar
FMAs, added last).  Here oracle-built graphs at dims 1 to 271 are searched on
the GPU and compared with the oracle bit for bit (ids, distances, counts).
The oracle's distances are pinned to the reference's own outputs at dims
1-200 by the distance fixtures, so this pins the GPU to the reference.

With ef >= N the beam search (include/hnsw.hpp:374-444) expands every node
it can reach; on these seeded graphs that yields the exact k smallest
distances for every query (true of the oracle's search on each case, also on
the few IP layers that are not strongly connected), so the returned
distances must equal brute force, whatever the ties at tiny dims do to the
ids.  Graphs of one and five nodes (fewer than k) return every node.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi
from tests.gpu_util import index_from_oracle

pytestmark = pytest.mark.gpu

DTYPES = {"u8": np.uint8, "i8": np.int8, "f32": np.float32, "f16": np.float16}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rows(kind, n, dim, rng):
    if kind == "u8":
        return rng.integers(0, 256, (n, dim)).astype(np.uint8)
    if kind == "i8":
        return rng.integers(-128, 128, (n, dim)).astype(np.int8)
    x = rng.normal(0, 1, (n, dim)).astype(np.float32)
    return x.astype(DTYPES[kind])


def oracle_case(kind, metric, n, nq, dim, M, efc, seed):
    rng = np.random.default_rng(seed)
    base, queries = rows(kind, n, dim, rng), rows(kind, nq, dim, rng)
    g = O.Graph(base, metric, M, efc, O.sample_levels(n, M, seed=seed))
    g.build(np.arange(n, dtype=np.int32))
    return base, queries, g


def check_exhaustive(ds, base, queries, metric, k, n):
    """Sorted returned distances equal the exact k smallest."""
    kk = min(k, n)
    want = np.array([np.sort([O.distance(q, b, metric) for b in base])[:kk] for q in queries], np.float32)
    np.testing.assert_array_equal(ds[:, :kk], want)


FLOAT_DIMS = [1, 3, 17, 24, 31, 100, 200, 271]  # tails 1, 3, 1, 8, 15, 4, 8, 15
INT_DIMS = [1, 3, 17, 100, 200, 512]


@pytest.mark.parametrize("kind,metric,dim",
                         [(k, m, d) for k in ("f32", "f16") for m in (O.L2, O.IP) for d in FLOAT_DIMS] +
                         [(k, m, d) for k in ("u8", "i8") for m in (O.L2, O.IP) for d in INT_DIMS])
def test_ragged_dims_match_oracle(kind, metric, dim):
    n, nq, M, efc = 400, 40, 8, 40
    say("ragged", kind, "L2" if metric == O.L2 else "IP", dim)
    base, queries, g = oracle_case(kind, metric, n, nq, dim, M, efc, seed=dim)
    idx = index_from_oracle(g)
    for k, ef in [(10, 40), (10, 400)]:
        ids, ds, cnt = idx.search(queries, k, ef)
        oi, od, on = g.search(queries, k, ef)
        tag = f"k{k} ef{ef}"
        np.testing.assert_array_equal(cnt, on, err_msg=tag)
        np.testing.assert_array_equal(ids, oi, err_msg=tag)
        np.testing.assert_array_equal(ds, od, err_msg=tag)
    check_exhaustive(ds, base, queries, metric, 10, n)
    idx.close()


@pytest.mark.parametrize("n", [1, 5])
@pytest.mark.parametrize("kind", ["u8", "f32"])
def test_graphs_smaller_than_k(kind, n):
    """Fewer nodes than k: every node comes back, in distance order, -1 / +inf
    after them; the same from an oracle-built graph and from a GPU build."""
    dim, M, efc, k = 24, 4, 16, 10
    base, queries, g = oracle_case(kind, O.L2, n, 8, dim, M, efc, seed=n)
    say("tiny graph", kind, n)
    loaded = index_from_oracle(g)
    built = abi.Index(DTYPES[kind], abi.L2, dim, M, efc, n)
    built.build(base, levels=O.sample_levels(n, M, seed=n), max_batch=4096)
    for idx in (loaded, built):
        ids, ds, cnt = idx.search(queries, k, 32)
        assert (cnt == n).all()
        assert (ids[:, n:] == -1).all() and np.isinf(ds[:, n:]).all()
        assert all(sorted(r[:n].tolist()) == list(range(n)) for r in ids)
        check_exhaustive(ds, base, queries, O.L2, k, n)
    oi, od, _ = g.search(queries, k, 32)
    np.testing.assert_array_equal(loaded.search(queries, k, 32)[0][:, :n], oi[:, :n])
    loaded.close()
    built.close()


@pytest.mark.parametrize("kind,metric,dim",
                         [(k, m, d) for k in ("f32", "f16") for m in (O.L2, O.IP) for d in FLOAT_DIMS] +
                         [(k, m, d) for k in ("u8", "i8") for m in (O.L2, O.IP) for d in INT_DIMS])
def test_fast_mode_ragged_dims(kind, metric, dim):
    """The fast mode on the same graphs: with ef >= N every width returns the
    exact k smallest distances; on float data, for every query without tied
    distances, width 1 returns the exact search's ids and distances at ef 40."""
    n, nq, M, efc = 400, 40, 8, 40
    say("fast ragged", kind, "L2" if metric == O.L2 else "IP", dim)
    base, queries, g = oracle_case(kind, metric, n, nq, dim, M, efc, seed=dim)
    idx = index_from_oracle(g)
    for width in (1, 3):
        idx.set_fast_tuning(width=width)
        ids, ds, cnt = idx.search_fast(queries, 10, 400)
        assert (cnt == 10).all()
        check_exhaustive(ds, base, queries, metric, 10, n)
        for row in ids:
            assert len(set(row.tolist())) == 10
    if kind in ("f32", "f16"):
        # width 1 equals the exact search for every query whose distances to
        # all rows differ in float32 (ties are ordered by id in the fast mode,
        # by heap layout in the reference; f16 at tiny dims ties often)
        d32 = np.array([[O.distance(q, b, metric) for b in base] for q in queries], np.float32)
        tie_free = np.array([len(np.unique(r)) == n for r in d32])
        if dim >= 17:
            assert tie_free.mean() > 0.9
        idx.set_fast_tuning(width=1)
        want = idx.search(queries, 10, 40)
        got = idx.search_fast(queries, 10, 40)
        say("  tie-free queries compared:", int(tie_free.sum()), "of", nq)
        np.testing.assert_array_equal(got[2][tie_free], want[2][tie_free])
        np.testing.assert_array_equal(got[0][tie_free], want[0][tie_free])
        np.testing.assert_array_equal(got[1][tie_free], want[1][tie_free])
    idx.close()


@pytest.mark.parametrize("n", [1, 5])
def test_fast_mode_graphs_smaller_than_k(n):
    dim, M, efc, k = 24, 4, 16, 10
    base, queries, g = oracle_case("f32", O.L2, n, 8, dim, M, efc, seed=n)
    idx = index_from_oracle(g)
    ids, ds, cnt = idx.search_fast(queries, k, 32)
    assert (cnt == n).all()
    assert (ids[:, n:] == -1).all() and np.isinf(ds[:, n:]).all()
    check_exhaustive(ds, base, queries, O.L2, k, n)
    idx.close()
