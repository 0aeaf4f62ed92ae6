"""GPU insertion (Insert / SelectHeuristic / AddNode / AddEdge) through the C-ABI.

* exact: batches of one node with the reference's row semantics reproduce the
  reference's own single-threaded build bit for bit (the golden graphs);
* batched: the default batched build's recall-vs-ef curve stays close to the
  CPU-built graph's.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi
from tests import golden_cases as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", ["u8_l2", "i8_ip", "f32_l2", "f16_l2"])
def test_exact_build_matches_reference(name, graph):
    """Batch-1 builds with reference rows equal the reference's build.  With
    graph, every batch's two phases replay HIP graphs updated in place
    (thousands of replays, host work and a sync between them)."""
    fx, p, base, queries = G.graph_case(name)
    say("exact build", name, "graph" if graph else "stream", p)
    g = G.oracle_graph(fx, p, base)  # equal to the reference's build (test_oracle_golden)
    idx = abi.Index(base.dtype, abi.L2 if p["metric"] == "l2" else abi.IP, p["dim"], p["M"], p["efc"], p["n"])
    idx.build(base, levels=fx["levels"], max_batch=1, exact_rows=True, graph=graph)
    info = idx.info()
    assert info["enter_level"] == int(fx["enter_level"])
    for l in range(g.nlayers):
        want = g.layer(l)
        got = idx.read_layer(l)
        n = want["next_id"]
        np.testing.assert_array_equal(got["ext"][:n], want["ext"][:n], err_msg=f"layer {l} ext")
        np.testing.assert_array_equal(got["deg"][:n], want["deg"][:n].astype(np.int32), err_msg=f"layer {l} deg")
        np.testing.assert_array_equal(got["heur"][:n], want["heur"][:n], err_msg=f"layer {l} heur")
        np.testing.assert_array_equal(got["adj"][:n], want["adj"][:n], err_msg=f"layer {l} adj")
        np.testing.assert_array_equal(got["dist"][:n], want["dist"][:n], err_msg=f"layer {l} dist")
    for k, ef in fx["searches"]:
        tag = f"k{k}_ef{ef}"
        ids, ds, cnt = idx.search(queries, int(k), int(ef))
        np.testing.assert_array_equal(ids, fx[f"ids_{tag}"], err_msg=tag)
        np.testing.assert_array_equal(ds, fx[f"dists_{tag}"], err_msg=tag)
    idx.close()


def recall(ids, gt, k):
    return np.mean([len(set(a[:k]) & set(b[:k])) / k for a, b in zip(ids, gt)])


@pytest.mark.parametrize("reference_rows", [True, False])
def test_batched_build_recall_close_to_cpu_graph(reference_rows):
    """Batched GPU insertion (batches up to 4096, concurrent AddEdge under
    per-node locks; two builds, mean recall) measured by its recall-vs-ef curve against the graph the
    reference's algorithm builds one node at a time.  Both builds depend on
    the insertion order and, on the GPU, on batching, so one pair of 20k-node
    graphs differs by about 1% at a single ef either way; the bound is on the
    curve's mean gap over the ef sweep.  With reference rows (rows may grow
    past D as include/hnsw_layer.hpp:441-447 lets them) the mean gap is held
    to 0.5% and every ef to 2%.  Rows capped at D hold fewer edges and give
    up some recall at equal ef (about 2-3% at ef=10): held to 3% / 4%."""
    from pickle_amd import datasets

    n, nq, dim, M, efc = 20000, 1000, 64, 12, 80
    base, queries = datasets.gaussian_mixture(n, nq, dim, 32, 21, kind="u8")
    levels = O.sample_levels(n, M, seed=21)
    say("batched build", n, dim, M, efc, "reference rows" if reference_rows else "capped rows")
    g = O.Graph(base, O.L2, M, efc, levels)
    g.build(np.arange(n, dtype=np.int32))
    gt, _ = O.brute_force(base, queries, 10, O.L2)
    # two batched builds: the order of concurrent AddEdge calls differs from
    # build to build, and their mean recall halves that variance
    # (the second runs its batches as HIP graphs, PK_BUILD_GRAPH)
    builds = []
    for b in range(2):
        idx = abi.Index(np.uint8, abi.L2, dim, M, efc, n)
        idx.build(base, levels=levels, max_batch=4096, exact_rows=reference_rows, graph=b == 1)
        if not reference_rows:
            for l in range(idx.info()["nlayers"]):
                L = idx.read_layer(l)
                assert (L["deg"] <= L["adj"].shape[1]).all(), "rows stay within D"
        builds.append(idx)
    cpu_idx = abi.Index(np.uint8, abi.L2, dim, M, efc, n)
    cpu_idx.set_base(base)
    for l in range(g.nlayers):
        L = g.layer(l)
        cpu_idx.load_layer(l, L["ext"], L["deg"].astype(np.int32), L["adj"], L["has_data"], heur=L["heur"])
    cpu_idx.set_enter_level(g.enter_level)
    gaps = []
    for ef in (10, 20, 40, 80):
        r_gpu = np.mean([recall(b.search(queries, 10, ef)[0], gt, 10) for b in builds])
        r_cpu = recall(cpu_idx.search(queries, 10, ef)[0], gt, 10)
        say(f"  ef={ef} recall gpu-built {r_gpu:.4f} cpu-built {r_cpu:.4f}")
        assert r_gpu >= r_cpu - (0.02 if reference_rows else 0.04), (ef, r_gpu, r_cpu)
        gaps.append(r_gpu - r_cpu)
    say(f"  mean gap {np.mean(gaps):+.4f}")
    assert np.mean(gaps) >= -(0.005 if reference_rows else 0.03), gaps
    for b in builds:
        b.close()
    cpu_idx.close()
