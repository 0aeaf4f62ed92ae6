"""On-disk graph format: the reference serializer's byte layout plus a v2 sidecar.

Main file, little-endian, exactly the per-layer layout of
include/serializer.hpp:27-55:

    u64 num_layers
    per layer:
        u64 graph_size      24 + 4N + 4N + 4ND + 4ND
        u64 num_nodes       N  (layer capacity; every slot is filled after a build)
        u64 max_degree      D  (2M at layer 0, M above)
        i32 ext[N]          external id of each internal id
        i32 deg[N]          degree
        i32 adj[N*D]        neighbour internal ids, padded with -1
        f32 dist[N*D]       neighbour distances, padded with FLT_MAX

Sidecar ``<path>.v2meta`` holds what pickle v2 keeps and that layout has no
room for (hnsw_layer.hpp:60-65, hnsw.hpp:133-139):

    char magic[8] = "PKLV2SC\\0"; u32 version = 1
    u32 dtype (common.hpp:21-26 codes), u32 metric (1 = L2, 2 = IP), u32 dim,
    u32 M, u32 efc, u64 n_total, i32 enter_level, u32 num_layers
    per layer: i32 ent_id, i32 next_id, i16 heuristic[N], u8 has_vector[N]
    u8 insert_level[n_total]

Vectors are not stored (the reference format stores none either): layer l's
copy of node i is the base row ext[i] when has_vector[i] is set and zeros
otherwise (a node that opened a new top level never had AddNode run there,
hnsw.hpp:283-297).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = b"PKLV2SC\x00"


@dataclass
class Layer:
    ext: np.ndarray          # int32 [N]
    deg: np.ndarray          # int32 [N]
    adj: np.ndarray          # int32 [N, D]
    dist: np.ndarray         # float32 [N, D]
    heur: np.ndarray         # int16 [N]
    has_vector: np.ndarray   # uint8 [N]
    ent_id: int = 0
    next_id: int = 0

    @property
    def n(self) -> int:
        return int(self.ext.shape[0])

    @property
    def max_degree(self) -> int:
        return int(self.adj.shape[1])


@dataclass
class GraphFile:
    dtype: int
    metric: int
    dim: int
    M: int
    efc: int
    n_total: int
    enter_level: int
    layers: list = field(default_factory=list)
    levels: np.ndarray | None = None


def read(path: str) -> GraphFile:
    with open(path, "rb") as f:
        buf = f.read()
    with open(path + ".v2meta", "rb") as f:
        meta = f.read()
    if meta[:8] != MAGIC:
        raise ValueError(f"{path}.v2meta: bad magic")
    (ver, dtype, metric, dim, M, efc) = struct.unpack_from("<6I", meta, 8)
    if ver != 1:
        raise ValueError(f"unsupported sidecar version {ver}")
    n_total, enter, nlm = struct.unpack_from("<QiI", meta, 32)
    mo = 32 + 16
    (nl,) = struct.unpack_from("<Q", buf, 0)
    if nl != nlm:
        raise ValueError("layer count differs between graph file and sidecar")
    off = 8
    layers = []
    for _ in range(nl):
        size, n, D = struct.unpack_from("<3Q", buf, off)
        if size != 24 + 8 * n + 8 * n * D:
            raise ValueError("graph_size field does not match the layer shape")
        off += 24
        ext = np.frombuffer(buf, np.int32, n, off).copy(); off += 4 * n
        deg = np.frombuffer(buf, np.int32, n, off).copy(); off += 4 * n
        adj = np.frombuffer(buf, np.int32, n * D, off).reshape(n, D).copy(); off += 4 * n * D
        dist = np.frombuffer(buf, np.float32, n * D, off).reshape(n, D).copy(); off += 4 * n * D
        ent_id, next_id = struct.unpack_from("<ii", meta, mo); mo += 8
        heur = np.frombuffer(meta, np.int16, n, mo).copy(); mo += 2 * n
        has = np.frombuffer(meta, np.uint8, n, mo).copy(); mo += n
        layers.append(Layer(ext, deg, adj, dist, heur, has, ent_id, next_id))
    levels = np.frombuffer(meta, np.uint8, n_total, mo).copy()
    return GraphFile(dtype, metric, dim, M, efc, n_total, enter, layers, levels)


def write(path: str, g: GraphFile) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(g.layers)))
        for L in g.layers:
            n, D = L.n, L.max_degree
            f.write(struct.pack("<3Q", 24 + 8 * n + 8 * n * D, n, D))
            f.write(np.ascontiguousarray(L.ext, np.int32).tobytes())
            f.write(np.ascontiguousarray(L.deg, np.int32).tobytes())
            f.write(np.ascontiguousarray(L.adj, np.int32).tobytes())
            f.write(np.ascontiguousarray(L.dist, np.float32).tobytes())
    with open(path + ".v2meta", "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<6I", 1, g.dtype, g.metric, g.dim, g.M, g.efc))
        f.write(struct.pack("<QiI", g.n_total, g.enter_level, len(g.layers)))
        for L in g.layers:
            f.write(struct.pack("<ii", L.ent_id, L.next_id))
            f.write(np.ascontiguousarray(L.heur, np.int16).tobytes())
            f.write(np.ascontiguousarray(L.has_vector, np.uint8).tobytes())
        lv = g.levels if g.levels is not None else np.zeros(g.n_total, np.uint8)
        f.write(np.ascontiguousarray(lv, np.uint8).tobytes())


def layer_vectors(g: GraphFile, level: int, base: np.ndarray) -> np.ndarray:
    """Layer `level`'s per-internal-id vector copy rebuilt from the base rows."""
    L = g.layers[level]
    out = np.zeros((L.n, base.shape[1]), dtype=base.dtype)
    sel = (L.has_vector != 0) & (L.ext >= 0)
    out[sel] = base[L.ext[sel]]
    return out
