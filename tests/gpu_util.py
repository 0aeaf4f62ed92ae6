"""Loading oracle-built graphs into the HIP index (shared by the GPU tests)."""
import numpy as np

from pickle_amd import abi


def index_from_oracle(g, capacity=None, slack=True):
    """An abi.Index holding exactly the oracle graph g (its layers, the slack
    row written past a full last row, and the per-layer vector flags)."""
    idx = abi.Index(g.base.dtype, g.metric, g.dim, g.M, g.efc, capacity or g.base.shape[0])
    idx.set_base(g.base)
    for l in range(g.nlayers):
        L = g.layer(l)
        cap = L["cap"]
        adj = L["adj"] if slack else L["adj"][:cap]
        idx.load_layer(l, L["ext"], L["deg"].astype(np.int32), adj, L["has_data"], heur=L["heur"],
                       dist=L["dist"][:cap])
    idx.set_enter_level(g.enter_level)
    return idx


def vector_flags(idx):
    """Per layer, which nodes hold their vector in that layer's copy (the
    graph file's sidecar, written by pk_index_save_file).  A node that
    opened a new top level never had AddNode run there (hnsw.hpp:283-297):
    its copy on those levels is zeros."""
    import os
    import tempfile

    from pickle_amd import serializer

    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "g.bin")
        idx.save_file(path)
        gf = serializer.read(path)
    return [L.has_vector.copy() for L in gf.layers]


def oracle_from_index(idx, base, metric, M, efc, levels):
    """An oracle graph holding the layers the GPU index holds (read back
    through pk_index_read_layer): rows of D ids (no slack past the last row;
    use with graphs whose rows stay within D), and each layer's vector copy
    as the index holds it (zeros for a node that opened a new top level)."""
    from oracle import oracle as O

    g = O.Graph(base, metric, M, efc, levels)
    info = idx.info()
    row_bytes = base.dtype.itemsize * base.shape[1]
    has_all = vector_flags(idx)
    for l in range(info["nlayers"]):
        L = idx.read_layer(l)
        n, D = L["adj"].shape
        adj = np.full((n + 1, D), -1, np.int32)
        adj[:n] = L["adj"]
        dist = np.full((n + 1, D), np.finfo(np.float32).max, np.float32)
        dist[:n] = L["dist"]
        has = has_all[l][:n].astype(np.uint8)
        data = np.ascontiguousarray(base[L["ext"]]).view(np.uint8).reshape(n, row_bytes).copy()
        data[has == 0] = 0
        g.set_layer(l, {"cap": n, "D": D, "next_id": n, "adj": adj, "dist": dist,
                        "deg": L["deg"].astype(np.int16), "heur": L["heur"], "ext": L["ext"],
                        "data": data, "has_data": has})
    g.set_enter_level(info["enter_level"])
    return g


def row_structure(idx, level, D_cap=None, overlapped_ok=False):
    """(rows with an id outside the layer or -1 within the degree, rows with
    a repeated id, rows longer than D_cap) of one layer.  With overlapped_ok,
    a repeat does not count in a row whose storage an earlier row reaches
    (a reference row longer than D continues into the next rows' slots,
    hnsw_layer.hpp:441-447, and a write there can repeat an id)."""
    L = idx.read_layer(level)
    n, D = L["adj"].shape
    ends = np.arange(n, dtype=np.int64) * D + L["deg"].astype(np.int64)  # flat end of every row
    bad_id = dup = too_long = 0
    for v in range(n):
        d = int(L["deg"][v])
        if D_cap is not None and d > D_cap:
            too_long += 1
        row = L["adj"][v, :min(d, D)]
        if (row < 0).any() or (row >= n).any():
            bad_id += 1
        if len(set(row.tolist())) != len(row):
            lo = max(0, v - 8)
            if not (overlapped_ok and (ends[lo:v] > v * D).any()):
                dup += 1
    return bad_id, dup, too_long
