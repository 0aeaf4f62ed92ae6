"""Precision form on the GPU: FP16 / per-dimension int8 traversal copies with
an exact FP32 re-rank (pk_search_rerank), checked against the oracle.

Pinning: the oracle graph is built on the FP32 vectors; the same graph is
stored again with the traversal copy's vectors (FP16 by round-to-nearest-even,
or the int8 quantizer restated in numpy), and the oracle's AnnSearch on it
with k = ef gives the candidates.  The expected answer is those candidates
ranked by the oracle's FP32 distance (the reference's summation order),
ties by candidate rank.

int8 traversal distances are exact integers, so ids and distances must match
bit for bit.  An FP16 copy traverses with its own FP32 summation order (16-byte
gathers), so its traversal distances agree with the reference order within
1e-5 relative: an id may differ from the expected one only where the two
distances at that rank agree within 1e-5 relative (a near-tie), and every
returned distance must be the exact FP32 distance (reference order) of the id
returned.  Both shapes run: d=64, M=12, and the 10M workload's d=128, M=32;
and both traversals (exact and fast mode) are re-ranked.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi, datasets, precision
from tests.gpu_util import index_from_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def quantize_i8(x, offset, scale):
    """numpy restatement of the device quantizer (FP32 arithmetic)."""
    inv = (np.float32(1.0) / scale.astype(np.float32)).astype(np.float32)
    u = (x.astype(np.float32) - offset.astype(np.float32)).astype(np.float32)
    u = (u * inv).astype(np.float32)
    return np.clip(np.rint(u), -127, 127).astype(np.int8)


def relayer(g32, conv, base_conv, dim):
    """The oracle graph g32 with its layer vectors replaced by conv(rows);
    rows without a vector stay zero, as on the device."""
    g = O.Graph(base_conv, g32.metric, g32.M, g32.efc, g32.levels)
    for l in range(g32.nlayers):
        L = dict(g32.layer(l))
        cap = int(L["cap"])
        rows32 = np.frombuffer(np.ascontiguousarray(L["data"]).tobytes(), np.float32).reshape(-1, dim)[:cap]
        conv_rows = conv(rows32)
        conv_rows[L["has_data"][:cap] == 0] = 0
        L["data"] = np.ascontiguousarray(conv_rows).view(np.uint8).reshape(-1)
        g.set_layer(l, L)
    g.set_enter_level(g32.enter_level)
    return g


def expected(gconv, qconv, queries, base, metric, k, ef):
    cand, _, cnt = gconv.search(qconv, ef, ef)
    ids = np.full((len(queries), k), -1, np.int32)
    ds = np.full((len(queries), k), np.inf, np.float32)
    for i in range(len(queries)):
        c = cand[i, : cnt[i]]
        d = np.array([O.distance(queries[i], base[j], metric) for j in c], np.float32)
        order = np.lexsort((np.arange(len(c)), d))[:k]
        ids[i, : len(order)] = c[order]
        ds[i, : len(order)] = d[order]
    return ids, ds


def near_ties_only(got, want, tag, rtol=1e-5):
    """Positions where the ids differ hold distances equal within rtol."""
    gi, gd = got
    wi, wd = want
    diff = gi != wi
    if diff.any():
        a, b = gd[diff].astype(np.float64), wd[diff].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        assert rel.max() <= rtol, (tag, int(diff.sum()), float(rel.max()))
    return int(diff.sum())


SHAPES = {"d64": (3000, 64, 64, 12, 60, 10, 40), "d128": (8000, 64, 128, 32, 100, 10, 64)}


@pytest.mark.parametrize("shape", ["d64", "d128"])
@pytest.mark.parametrize("kind,metric", [("f16", O.L2), ("i8", O.L2), ("f16", O.IP), ("i8", O.IP)])
def test_rerank_matches_oracle(kind, metric, shape):
    n, nq, dim, M, efc, k, ef = SHAPES[shape]
    if metric == O.L2:
        base, queries = datasets.gaussian_mixture(n, nq, dim, 16, 31, kind="f32")
    else:
        base, queries = datasets.vmf_mixture(n, nq, dim, 16, 31)
    say("precision", kind, "L2" if metric == O.L2 else "IP", n, dim)
    levels = O.sample_levels(n, M, seed=31)
    g32 = O.Graph(base, metric, M, efc, levels)
    g32.build(np.arange(n, dtype=np.int32))
    if kind == "f16":
        conv = lambda x: x.astype(np.float16)
        base_c, q_c = base.astype(np.float16), queries.astype(np.float16)
    else:
        off, sc = precision.int8_quantizer(base, metric)
        conv = lambda x: quantize_i8(x, off, sc)
        base_c, q_c = quantize_i8(base, off, sc), quantize_i8(queries, off, sc)
    gconv = relayer(g32, conv, base_c, dim)
    want_ids, want_d = expected(gconv, q_c, queries, base, metric, k, ef)

    src = index_from_oracle(g32, slack=False)  # the re-layered oracle graphs hold no slack row
    idx = precision.traversal_index(src, base, kind)
    ids, ds, cnt = idx.search_rerank(queries, k, ef)
    assert (cnt == k).all()
    t_ids, t_d, _ = idx.search(q_c, 10, ef)  # the traversal copy alone, through the ordinary search
    o_ids, o_d, _ = gconv.search(q_c, 10, ef)
    if kind == "i8":
        np.testing.assert_array_equal(ids, want_ids)
        np.testing.assert_array_equal(ds, want_d)
        np.testing.assert_array_equal(t_ids, o_ids)
        np.testing.assert_array_equal(t_d, o_d)
    else:
        n_diff = near_ties_only((ids, ds), (want_ids, want_d), "re-ranked")
        exact = np.array([[O.distance(queries[i], base[j], metric) for j in row] for i, row in enumerate(ids)],
                         np.float32)
        np.testing.assert_array_equal(ds, exact)  # the re-rank's FP32 values, reference order
        ref_t = np.array([[O.distance(q_c[i], base_c[j], metric) for j in row] for i, row in enumerate(t_ids)],
                         np.float32)
        np.testing.assert_allclose(t_d, ref_t, rtol=1e-5, atol=1e-6)
        near_ties_only((t_ids, ref_t), (o_ids, o_d), "traversal")
        say("  ids differing at near-ties:", n_diff, "of", ids.size)
    # the fast traversal, re-ranked the same way: recall-equivalent
    gt, _ = O.brute_force(base, queries, k, metric)
    rec = lambda a: np.mean([len(set(x) & set(y)) / k for x, y in zip(a, gt)])
    f_ids = idx.search_rerank_fast(queries, k, ef)[0]
    assert rec(f_ids) >= rec(ids) - 0.01, (rec(f_ids), rec(ids))
    idx.close()
    src.close()


def test_rerank_recall_and_errors():
    n, nq, dim, M, efc = 20000, 300, 96, 16, 100
    base, queries = datasets.gaussian_mixture(n, nq, dim, 32, 33, kind="f32")
    gt, _ = O.brute_force(base, queries, 10, O.L2)
    src = abi.Index(np.float32, abi.L2, dim, M, efc, n)
    src.build(base, seed=33)
    rec = lambda ids: np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)])
    r32 = rec(src.search(queries, 10, 64)[0])
    for kind in ("f16", "i8"):
        idx = precision.traversal_index(src, base, kind)
        r = rec(idx.search_rerank(queries, 10, 64)[0])
        say(f"  recall@10 ef=64: fp32 traversal {r32:.4f}, {kind} traversal + fp32 re-rank {r:.4f}")
        assert r >= r32 - 0.01
        with pytest.raises(abi.PickleError):
            idx.search_rerank(queries, 10, 5)  # ef < k
        # 256-slot visited tables, forced from the host: a search at ef 300
        # keeps 300 distinct ids, so every traversal fills its table.  The
        # device call reports it; the host call runs those queries again with
        # larger tables and returns what the default tables give.
        want = idx.search_rerank(queries, 10, 300)
        idx.set_tuning(0, 8, 0)
        nq = queries.shape[0]
        dq = abi.DeviceBuffer.from_array(np.ascontiguousarray(queries, np.float32))
        d_ids, d_d, d_c = abi.DeviceBuffer(nq * 40), abi.DeviceBuffer(nq * 40), abi.DeviceBuffer(nq * 4)
        idx.search_rerank_device(dq, nq, 10, 300, 10, d_ids, d_d, d_c)
        idx.sync()
        assert idx.device_error() == 1
        assert (d_c.to_array(np.int32, (nq,)) == -1).all()
        for a, b in zip(idx.search_rerank(queries, 10, 300), want):
            np.testing.assert_array_equal(a, b)
        for buf in (dq, d_ids, d_d, d_c):
            buf.free()
        idx.set_tuning(0, 0, 0)
        assert rec(idx.search_rerank(queries, 10, 64)[0]) == r  # the tables were left empty
        idx.close()
    with pytest.raises(abi.PickleError):
        src.set_full_rows(base)  # an FP32 index has no separate traversal copy
    src.close()
