"""Helpers shared by the tests that read tests/golden/*.npz."""
import hashlib
import json
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRAPH_CASES = ["u8_l2", "i8_l2", "u8_ip", "i8_ip", "f32_l2", "f32_ip", "f16_l2", "f16_ip"]


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def graph_case(name):
    """(fixture, params, base, queries) with the vectors regenerated and checked."""
    from tests.golden.make_golden import make_vectors  # generator definitions live with the fixtures

    fx = load(f"graph_{name}.npz")
    p = json.loads(str(fx["params"]))
    base, queries = make_vectors(tuple(p["gen"]), p["n"], p["nq"], p["dim"])
    assert sha(base) == str(fx["base_sha"]), "generator stream changed: base vectors differ"
    assert sha(queries) == str(fx["queries_sha"]), "generator stream changed: queries differ"
    return fx, p, base, queries


def metric_code(m):
    return O.L2 if m == "l2" else O.IP


def oracle_graph(fx, p, base):
    g = O.Graph(base, metric_code(p["metric"]), p["M"], p["efc"], fx["levels"])
    g.build(np.arange(p["n"], dtype=np.int32))
    return g


def oracle_graph_digest(g):
    h = hashlib.sha256()
    for l in range(g.nlayers):
        L = g.layer(l)
        cap = L["cap"]
        for a in (L["ext"], L["deg"].astype(np.int32), L["heur"], L["adj"][:cap], L["dist"][:cap]):
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
