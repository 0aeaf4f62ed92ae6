"""The serving loop (pk_server_*) on the GPU.

Requests of mixed size, k and ef are submitted from several threads at once;
each result must equal the same queries searched alone (the search is exact,
so the mix cannot change a result).  Streaming inserts go through the loop
and are visible to the searches admitted after them."""
import sys
import threading

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi, datasets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_requests_alone_equal_mixed_and_inserts_visible():
    n0, extra, dim, M, efc = 20000, 2000, 64, 12, 80
    base, queries = datasets.gaussian_mixture(n0 + extra, 3000, dim, 32, 41, kind="u8")
    idx = abi.Index(np.uint8, abi.L2, dim, M, efc, n0 + extra)
    levels = abi.sample_levels(n0 + extra, M, 41)
    idx.set_base(base[:n0])
    idx.init_build(levels)
    idx.insert(np.arange(n0, dtype=np.int32), max_batch=4096)
    rng = np.random.default_rng(5)
    reqs, at = [], 0
    while at < len(queries):
        m = int(min(len(queries) - at, np.exp(rng.uniform(0, np.log(400)))))
        reqs.append((at, m, int(rng.choice([1, 10, 32])), int(rng.integers(32, 97))))
        at += m
    say("requests", len(reqs))
    alone = [idx.search(queries[s:s + m], k, ef) for s, m, k, ef in reqs]

    srv = abi.Server(idx, slots=4, slot_queries=512, max_k=32, max_ef=128)
    results = [None] * len(reqs)

    def client(part):
        tickets = [(i, srv.submit(queries[reqs[i][0]:reqs[i][0] + reqs[i][1]], reqs[i][2], reqs[i][3]))
                   for i in part]
        for i, t in tickets:
            results[i] = srv.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    st = srv.stats()
    say("stats", st)
    for i, (s, m, k, ef) in enumerate(reqs):
        ids, ds, cnt, lat, que = results[i]
        np.testing.assert_array_equal(ids, alone[i][0][:, :k], err_msg=f"request {i}")
        np.testing.assert_array_equal(ds, alone[i][1][:, :k], err_msg=f"request {i}")
        np.testing.assert_array_equal(cnt, alone[i][2], err_msg=f"request {i}")
        assert lat >= que >= 0
    assert st["queries"] == len(queries) and st["launches"] < len(reqs)
    assert st["max_concurrent"] >= 2

    # streaming insert, then the new rows find themselves
    new_ids = np.arange(n0, n0 + extra, dtype=np.int32)
    t_ins = srv.insert(new_ids, base[n0:])
    t_mid = srv.submit(queries[:300], 10, 64)  # may run between the insert's steps
    assert srv.wait(t_ins)[0] is None
    assert srv.wait(t_mid)[2].min() == 10
    t_q = srv.submit(base[n0:n0 + 500], 1, 64)  # submitted after the insert finished
    ids, ds, _, _, _ = srv.wait(t_q)
    found = float(np.mean(ids[:, 0] == new_ids[:500]))
    say("inserted rows found as their own nearest neighbour:", found)
    assert found >= 0.95
    srv.close()
    # the loop's post-insert answers equal the eager search of the same graph
    np.testing.assert_array_equal(ids, idx.search(base[n0:n0 + 500], 1, 64)[0])
    idx.close()


def test_request_that_overflows_scratch_reruns_alone():
    """A query that fills its visited table (forced from the host with
    2^8-slot tables) ends early with count -1; the loop runs the failed rows
    again with tables twice the size until they fit, and the request
    completes with the exact results, while the other requests, in the same
    launch or not, return theirs.  On the 3000-node golden graph a query at
    ef 1 marks at most 159 ids per level, one at ef 300 at least 485."""
    from tests import golden_cases as G
    from tests.gpu_util import index_from_oracle

    fx, p, base, queries = G.graph_case("f32_l2")
    g = G.oracle_graph(fx, p, base)
    idx = index_from_oracle(g)
    idx.set_tuning(0, 8, 0)
    srv = abi.Server(idx, slots=1, slot_queries=128, max_k=10, max_ef=300)
    # the handle's own search still runs while a server holds the index (its
    # scratch is first allocated here)
    ids, ds, _ = idx.search(queries[:10], 10, 40)
    oi, od, _ = g.search(queries[:10], 10, 40)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)
    t_ok = srv.submit(queries[:50], 1, 1)
    t_big = srv.submit(queries[50:], 10, 300)
    t_ok2 = srv.submit(queries[50:], 1, 1)
    say("three requests on 256-slot visited tables")
    for t, sl, k, ef in ((t_big, slice(50, None), 10, 300), (t_ok, slice(0, 50), 1, 1), (t_ok2, slice(50, None), 1, 1)):
        ids, ds, cnt, _, _ = srv.wait(t)
        oi, od, on = g.search(queries[sl], k, ef)
        np.testing.assert_array_equal(ids, oi, err_msg=f"k{k} ef{ef}")
        np.testing.assert_array_equal(ds, od, err_msg=f"k{k} ef{ef}")
        np.testing.assert_array_equal(cnt, on, err_msg=f"k{k} ef{ef}")
    st = srv.stats()
    say("stats", st)
    assert st["queries"] == len(queries[:50]) + 2 * len(queries[50:])
    assert st["launches"] >= 3  # the failed rows ran again with 2^9, 2^10, ... slots
    with pytest.raises(abi.PickleError) as e:
        srv.wait(t_big)  # released
    assert e.value.code == -1
    srv.close()
    idx.close()


def test_fast_mode_requests_alone_equal_mixed():
    """PK_SERVE_FAST: launches run the fast mode with per-query k and ef and
    the launch's largest ef; each request equals the same queries searched
    alone with pk_search_fast_device (a query's result depends only on the
    graph, the query, k, ef and the handle's fast tuning)."""
    n, dim, M, efc = 20000, 64, 12, 80
    base, queries = datasets.gaussian_mixture(n, 2000, dim, 32, 43, kind="u8")
    idx = abi.Index(np.uint8, abi.L2, dim, M, efc, n)
    idx.build(base, max_batch=4096, seed=43)
    rng = np.random.default_rng(7)
    reqs, at = [], 0
    while at < len(queries):
        m = int(min(len(queries) - at, np.exp(rng.uniform(0, np.log(300)))))
        reqs.append((at, m, int(rng.choice([1, 10, 32])), int(rng.integers(32, 300))))
        at += m
    say("requests", len(reqs))
    alone = [idx.search_fast(queries[s:s + m], k, ef) for s, m, k, ef in reqs]
    srv = abi.Server(idx, slots=4, slot_queries=512, max_k=32, max_ef=300, fast=True)
    results = [None] * len(reqs)

    def client(part):
        tickets = [(i, srv.submit(queries[reqs[i][0]:reqs[i][0] + reqs[i][1]], reqs[i][2], reqs[i][3]))
                   for i in part]
        for i, t in tickets:
            results[i] = srv.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    st = srv.stats()
    say("stats", st)
    srv.close()
    for i, (s, m, k, ef) in enumerate(reqs):
        ids, ds, cnt, lat, que = results[i]
        np.testing.assert_array_equal(cnt, alone[i][2], err_msg=f"request {i}")
        np.testing.assert_array_equal(ids, alone[i][0][:, :k], err_msg=f"request {i}")
        np.testing.assert_array_equal(ds, alone[i][1][:, :k], err_msg=f"request {i}")
    assert st["queries"] == len(queries) and st["launches"] < len(reqs)
    with pytest.raises(abi.PickleError):
        abi.Server(idx, max_k=10, max_ef=600, fast=True)  # the fast mode's ef limit
    idx.close()
