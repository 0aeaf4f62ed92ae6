"""HIP search (through the C-ABI) against the reference's own results and the oracle.

Bit-exact ids, distances, result counts and work counters on every golden
graph (the reference's AnnSearch output), for all four dtypes and both metrics.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi, serializer
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


@pytest.mark.parametrize("name", G.GRAPH_CASES)
def test_search_matches_reference_bit_exact(graphs, name):
    fx, p, base, queries, g = graphs(name)
    say("case", name, p["dtype"], p["metric"], "n", p["n"], "dim", p["dim"])
    idx = index_from_oracle(g)
    for k, ef in fx["searches"]:
        tag = f"k{k}_ef{ef}"
        say("  search", tag)
        ids, ds, cnt, ctr = idx.search(queries, int(k), int(ef), counters=True)
        np.testing.assert_array_equal(cnt, fx[f"n_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"], err_msg=tag)
        want = fx[f"counters_{tag}"]  # hops, hops0, dists, dists0, neighbors
        got = [int(ctr["hops"].sum()), int(ctr["hops"][:, 0].sum()), int(ctr["dists"].sum()),
               int(ctr["dists"][:, 0].sum()), int(ctr["neighbors"].sum())]
        assert got == want.tolist(), tag
    idx.close()


def test_uneven_k_and_ef_per_query(graphs):
    fx, p, base, queries, g = graphs("u8_l2")
    rng = np.random.default_rng(7)
    nq = queries.shape[0]
    ks = rng.choice([1, 10, 100], nq).astype(np.int32)
    efs = np.exp(rng.uniform(np.log(16), np.log(256), nq)).astype(np.int32)
    idx = index_from_oracle(g)
    ids, ds, cnt = idx.search(queries, ks, efs, k_stride=100)
    for i in range(nq):
        oi, od, on = g.search(queries[i], int(ks[i]), int(efs[i]))
        assert cnt[i] == on[0]
        np.testing.assert_array_equal(ids[i, : ks[i]], oi[0])
        np.testing.assert_array_equal(ds[i, : ks[i]], od[0])
        assert (ids[i, ks[i]:] == -1).all()
    idx.close()


def test_spill_paths_exact(graphs):
    """A tiny LDS part forces the candidate heap into its device-memory
    spill, and a small visited table forces long probe chains; results stay
    identical."""
    fx, p, base, queries, g = graphs("f32_l2")
    idx = index_from_oracle(g)
    idx.set_tuning(ccap=8, vbits=12)
    for k, ef in [(10, 80), (100, 128)]:
        tag = f"k{k}_ef{ef}"
        ids, ds, cnt = idx.search(queries, k, ef)
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"])
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"])
    idx.close()


def test_full_visited_table_is_reported_rerun_and_left_clean(graphs):
    """A visited table too small for the search, forced from the host with
    2^8 slots, ends the query early: the device call reports flag 1 and
    count -1.  On this 3000-node graph a query at ef 300 marks at least 485
    ids on layer 0 (the oracle's work counters), one at ef 1 at most 159 on
    any layer, so ef 1 still fits and stays exact.  The host call pk_search
    searches the failed queries again with larger tables and returns the
    reference's results.  Each level empties its table also when it fails: a
    search with larger tables laid over the same memory afterwards gives the
    reference's results again."""
    fx, p, base, queries, g = graphs("f32_l2")
    idx = index_from_oracle(g)
    idx.set_tuning(0, 12, 0)
    for k, ef in [(10, 80), (100, 128)]:
        ids, ds, _ = idx.search(queries, k, ef)
        np.testing.assert_array_equal(ids, fx[f"ids_k{k}_ef{ef}"])
    idx.set_tuning(0, 8, 0)  # the same allocation, 256-slot tables packed at its start
    say("visited table of 256 slots, ef 300")
    nq = queries.shape[0]
    dq = abi.DeviceBuffer.from_array(np.ascontiguousarray(queries))
    d_ids, d_d, d_c = abi.DeviceBuffer(nq * 10 * 4), abi.DeviceBuffer(nq * 10 * 4), abi.DeviceBuffer(nq * 4)
    idx.search_device(dq, nq, 10, 300, 10, d_ids, d_d, d_c)
    idx.sync()
    assert idx.device_error() == 1
    assert (d_c.to_array(np.int32, (nq,)) == -1).all()
    # failed queries rerun with 2^9, 2^10, ... slots; their counters come back too
    ids, ds, cnt, ctr = idx.search(queries, 10, 300, counters=True)
    oi, od, on, octr = g.search(queries, 10, 300, counters=True)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)
    np.testing.assert_array_equal(cnt, on)
    for name in ("hops", "neighbors", "dists"):
        assert int(ctr[name][:, 0].sum()) == octr[name][0], name  # layer 0
        assert int(ctr[name].sum()) == sum(octr[name]), name
    for b in (dq, d_ids, d_d, d_c):
        b.free()
    ids, ds, cnt = idx.search(queries, 1, 1)
    oi, od, on = g.search(queries, 1, 1)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)
    np.testing.assert_array_equal(cnt, on)
    assert idx.device_error() == 0
    idx.set_tuning(0, 12, 0)
    for k, ef in [(10, 80), (100, 128)]:
        tag = f"k{k}_ef{ef}"
        ids, ds, cnt = idx.search(queries, k, ef)
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"], err_msg=tag)
    idx.close()


def test_graph_file_round_trip(graphs, tmp_path):
    """pk_index_load_file / save_file on the serializer layout + sidecar."""
    fx, p, base, queries, g = graphs("i8_l2")
    layers = []
    for l in range(g.nlayers):
        L = g.layer(l)
        cap = L["cap"]
        layers.append(serializer.Layer(L["ext"], L["deg"].astype(np.int32), L["adj"][:cap], L["dist"][:cap],
                                       L["heur"], L["has_data"], 0, L["next_id"]))
    gf = serializer.GraphFile(dtype=abi.INT8, metric=abi.L2, dim=p["dim"], M=p["M"], efc=p["efc"],
                              n_total=p["n"], enter_level=g.enter_level, layers=layers, levels=fx["levels"])
    path = str(tmp_path / "g.bin")
    serializer.write(path, gf)
    idx = abi.Index(np.int8, abi.L2, p["dim"], p["M"], p["efc"], p["n"])
    idx.set_base(base)
    idx.load_file(path)
    ids, ds, cnt = idx.search(queries, 10, 40)
    np.testing.assert_array_equal(ids, fx["ids_k10_ef40"])
    path2 = str(tmp_path / "g2.bin")
    idx.save_file(path2)
    g2 = serializer.read(path2)
    for a, b in zip(gf.layers, g2.layers):
        np.testing.assert_array_equal(a.adj, b.adj)
        np.testing.assert_array_equal(a.deg, b.deg)
        np.testing.assert_array_equal(a.heur, b.heur)
        np.testing.assert_array_equal(a.dist, b.dist)
    idx.close()

    # corrupt files are refused: a first-layer header whose node count makes
    # the size arithmetic wrap (2^62 nodes of degree 1 "fit" in 24 bytes),
    # and a file cut short
    raw = open(path, "rb").read()
    wrap = raw[:8] + np.array([24, 1 << 62, 1], np.uint64).tobytes() + raw[32:]
    for name, data in (("wrap.bin", wrap), ("short.bin", raw[: len(raw) // 2])):
        bad = str(tmp_path / name)
        open(bad, "wb").write(data)
        open(bad + ".v2meta", "wb").write(open(path + ".v2meta", "rb").read())
        idx = abi.Index(np.int8, abi.L2, p["dim"], p["M"], p["efc"], p["n"])
        idx.set_base(base)
        with pytest.raises(abi.PickleError) as e:
            idx.load_file(bad)
        assert e.value.code == -5, name  # PK_ERR_IO
        idx.close()


def test_duplicate_ids_in_a_row_follow_first_occurrence():
    """A row that names an id twice (possible after the reference's write past
    a full row) is scanned in order: the second copy is already visited."""
    rng = np.random.default_rng(3)
    base = rng.integers(0, 256, (400, 32)).astype(np.uint8)
    q = rng.integers(0, 256, (30, 32)).astype(np.uint8)
    lv = np.zeros(400, np.uint8)
    g = O.Graph(base, O.L2, 8, 40, lv)
    g.build(np.arange(400, dtype=np.int32))
    L = g.layer(0)
    adj = L["adj"].copy()
    deg = L["deg"].astype(np.int32).copy()
    for v in range(0, 400, 3):  # repeat an earlier neighbour at the row end
        d = deg[v]
        if 2 <= d < L["D"]:
            adj[v, d] = adj[v, int(rng.integers(0, d))]
            deg[v] = d + 1
    L2 = dict(L)
    L2["adj"] = adj
    L2["deg"] = deg.astype(np.int16)
    g.set_layer(0, L2)
    want = g.search(q, 10, 32, counters=True)
    idx = abi.Index(np.uint8, abi.L2, 32, 8, 40, 400)
    idx.set_base(base)
    idx.load_layer(0, L["ext"], deg, adj, L["has_data"], heur=L["heur"])
    idx.set_enter_level(0)
    ids, ds, cnt, ctr = idx.search(q, 10, 32, counters=True)
    np.testing.assert_array_equal(ids, want[0])
    np.testing.assert_array_equal(ds, want[1])
    assert int(ctr["dists"].sum()) == sum(want[3]["dists"])
    idx.close()


def test_bad_arguments_are_errors():
    idx = abi.Index(np.uint8, abi.L2, 16, 8, 10, 10)
    with pytest.raises(abi.PickleError):
        idx.search(np.zeros((2, 16), np.uint8), 10, 10)  # no graph loaded yet
    base = np.arange(160, dtype=np.uint8).reshape(10, 16)
    idx.set_base(base)
    idx.load_layer(0, np.arange(10), np.zeros(10, np.int32), -np.ones((10, 16), np.int32), np.ones(10, np.uint8))
    idx.set_enter_level(0)
    ids, ds, cnt = idx.search(base[:3], 1, 4)  # only the entry node is reachable
    assert (cnt == 1).all() and (ids[:, 0] == 0).all()
    with pytest.raises(abi.PickleError):
        idx.search(base[:3], np.array([1, 2, 3], np.int32), 4, k_stride=2)
    idx.close()


def test_brute_force_matches_fp64_oracle():
    rng = np.random.default_rng(11)
    for dt, metric in [(np.uint8, abi.L2), (np.int8, abi.IP), (np.float32, abi.L2), (np.float16, abi.IP)]:
        say("brute force", np.dtype(dt).name, metric)
        if np.dtype(dt).kind in "ui":
            info = np.iinfo(dt)
            base = rng.integers(info.min, info.max + 1, (3000, 100)).astype(dt)
            q = rng.integers(info.min, info.max + 1, (70, 100)).astype(dt)
        else:
            base = rng.normal(size=(3000, 100)).astype(dt)
            q = rng.normal(size=(70, 100)).astype(dt)
        idx = abi.Index(dt, metric, 100, 8, 10, 3000)
        idx.set_base(base)
        ids, ds = idx.brute_force(q)
        gt_ids, gt_d = O.brute_force(base, q, 10, O.L2 if metric == abi.L2 else O.IP)
        if np.dtype(dt).kind in "ui":
            np.testing.assert_array_equal(ids[:, :10], gt_ids)
            np.testing.assert_array_equal(ds[:, :10], gt_d.astype(np.float32))
        else:
            # FP64 re-rank of the 32 candidates gives the exact top 10
            A, Q = base.astype(np.float64), q.astype(np.float64)
            for i in range(len(q)):
                cand = ids[i]
                d64 = ((A[cand] - Q[i]) ** 2).sum(1) if metric == abi.L2 else -(A[cand] @ Q[i])
                order = np.lexsort((cand, d64))[:10]
                np.testing.assert_array_equal(cand[order], gt_ids[i])
        idx.close()
