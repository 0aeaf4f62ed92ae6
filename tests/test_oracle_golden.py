"""The oracle against the reference's own outputs (tests/golden, written by
tests/golden/make_golden.py from the compiled reference)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as G


def test_distance_known_answers():
    fx = G.load("dist_kat.npz")
    keys = [k[2:] for k in fx.files if k.startswith("d_")]
    assert len(keys) == 4 * 8 * 2
    for key in keys:
        dt, dim, metric = key.split("_")
        a, b, want = fx[f"a_{key}"], fx[f"b_{key}"], fx[f"d_{key}"]
        mc = O.L2 if metric == "l2" else O.IP
        got = np.array([O.distance(a[i], b[i], mc) for i in range(len(a))], np.float32)
        np.testing.assert_array_equal(got, want, err_msg=key)


def test_distance_closed_form():
    """Integer distances are exact sums; float ones within 1e-6 of FP64."""
    rng = np.random.default_rng(5)
    for dt in (np.uint8, np.int8, np.float16, np.float32):
        for dim in (7, 96, 128):
            if np.dtype(dt).kind in "ui":
                info = np.iinfo(dt)
                a = rng.integers(info.min, info.max + 1, (20, dim)).astype(dt)
                b = rng.integers(info.min, info.max + 1, (20, dim)).astype(dt)
            else:
                a = rng.normal(0, 3, (20, dim)).astype(dt)
                b = rng.normal(0, 3, (20, dim)).astype(dt)
            A, B = a.astype(np.float64), b.astype(np.float64)
            l2 = ((A - B) ** 2).sum(1)
            ip = -(A * B).sum(1)
            got_l2 = np.array([O.distance(a[i], b[i], O.L2) for i in range(20)])
            got_ip = np.array([O.distance(a[i], b[i], O.IP) for i in range(20)])
            if np.dtype(dt).kind in "ui":
                np.testing.assert_array_equal(got_l2, l2.astype(np.float32))
                np.testing.assert_array_equal(got_ip, ip.astype(np.float32))
            else:
                np.testing.assert_allclose(got_l2, l2, rtol=1e-6)
                np.testing.assert_allclose(got_ip, ip, rtol=1e-5, atol=1e-4 * np.abs(A * B).sum(1).max())


def test_heap_known_answers():
    fx = G.load("heap_kat.npz")
    for case in range(12):
        ops = fx[f"ops_{case}"]
        vals = fx[f"vals_{case}"]
        is_max = int(fx[f"is_max_{case}"])
        k = int(fx[f"final_k_{case}"])
        ids = np.arange(len(ops), dtype=np.int32)
        push_d = vals[ops == 1]
        push_i = ids[ops == 1]
        got = O.heap_script(bool(is_max), ops, push_d, push_i, final_k=k)
        want = list(zip(fx[f"pop_dist_{case}"].tolist(), fx[f"pop_id_{case}"].tolist()))
        assert [(float(d), int(i)) for d, i in got] == [(float(d), int(i)) for d, i in want], case


def test_heap_property_and_sorted_topk():
    """Closed-form checks that do not depend on tie order."""
    rng = np.random.default_rng(9)
    vals = rng.integers(0, 1000, 300).astype(np.float32)
    ops = np.ones(300, np.int8)
    got = O.heap_script(False, np.concatenate([ops, np.zeros(300, np.int8)]), vals, np.arange(300))
    assert [d for d, _ in got] == sorted(vals.tolist())
    got = O.heap_script(True, ops, vals, np.arange(300), final_k=25)
    assert [d for d, _ in got] == sorted(vals.tolist())[:25]


def test_layer_add_edge_known_answers():
    fx = G.load("layer_kat.npz")
    overflowed = 0
    for case in range(4):
        D = int(fx[f"D_{case}"])
        vecs = fx[f"vecs_{case}"]
        cap = vecs.shape[0]
        base = np.zeros((cap + 1, vecs.shape[1]), np.float32)
        base[:cap] = vecs
        g = O.Graph(base, O.L2, D // 2, 10, np.zeros(cap + 1, np.uint8))
        for (kind, a, b), d in zip(fx[f"ops_{case}"], fx[f"opd_{case}"]):
            if kind == 0:
                g.new_in_id(0, int(a))
            elif kind == 1:
                g.add_node(0, int(a), [], [], base[a])
            else:
                g.add_edge(0, int(a), int(b), float(d))
        L = g.layer(0)
        rows = cap + 1
        np.testing.assert_array_equal(L["adj"][:rows], fx[f"adj_{case}"])
        np.testing.assert_array_equal(L["dist"][:rows], fx[f"dist_{case}"])
        np.testing.assert_array_equal(L["deg"][:rows], fx[f"deg_{case}"])
        np.testing.assert_array_equal(L["heur"][:rows], fx[f"heur_{case}"])
        overflowed += int(fx[f"deg_{case}"].max() > D)
    assert overflowed >= 1, "fixtures must exercise the write past a full row"


@pytest.mark.parametrize("name", G.GRAPH_CASES)
def test_graph_build_and_search_match_reference(name):
    fx, p, base, queries = G.graph_case(name)
    g = G.oracle_graph(fx, p, base)
    assert g.nlayers == int(fx["nlayers"])
    assert g.enter_level == int(fx["enter_level"])
    assert G.oracle_graph_digest(g) == str(fx["graph_sha"])
    L0 = g.layer(0)
    np.testing.assert_array_equal(L0["adj"][: L0["cap"]], fx["l0_adj"])
    for k, ef in fx["searches"]:
        tag = f"k{k}_ef{ef}"
        ids, ds, ns, ctr = g.search(queries, int(k), int(ef), counters=True)
        np.testing.assert_array_equal(ns, fx[f"n_{tag}"])
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"], err_msg=tag)
        want = fx[f"counters_{tag}"]
        assert [sum(ctr["hops"]), ctr["hops"][0], sum(ctr["dists"]), ctr["dists"][0],
                sum(ctr["neighbors"])] == want.tolist(), tag


def test_levels_closed_form():
    """P(level >= l) = M^(-l(l+1)/2) for the fresh-draw-per-level sampler."""
    M = 4
    cum = O.cumulative_probability(M)
    np.testing.assert_allclose(cum[:3], [1 - M ** -1, 1 - M ** -2, 1 - M ** -3], rtol=1e-12)
    n = 400_000
    lv = O.sample_levels(n, M, seed=3)
    for l, expect in [(1, M ** -1), (2, M ** -3), (3, M ** -6)]:
        frac = np.mean(lv >= l)
        sd = np.sqrt(expect * (1 - expect) / n)
        assert abs(frac - expect) < 5 * sd + 1e-9, (l, frac, expect)


def test_search_brute_force_when_ef_covers_graph():
    """With ef >= N on a connected layer the beam search is exhaustive."""
    rng = np.random.default_rng(4)
    base = rng.integers(0, 256, (300, 16)).astype(np.uint8)
    q = rng.integers(0, 256, (10, 16)).astype(np.uint8)
    lv = O.sample_levels(300, 8, seed=1)
    g = O.Graph(base, O.L2, 8, 64, lv)
    g.build(np.arange(300, dtype=np.int32))
    ids, ds, _ = g.search(q, 10, 400)
    gt_ids, gt_d = O.brute_force(base, q, 10, O.L2)
    np.testing.assert_array_equal(ds, gt_d.astype(np.float32))
