"""ctypes wrapper for liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of pickle v2's HNSW path (oracle.c cites
the reference file:line for each function).  Only tests/, the smoke check
and bench.py's cpu_baseline leg import this module; the product
(pickle_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "pickle_ref")
REF_BIN_GENERIC = os.path.join(HERE, "_ref", "pickle_ref_generic")

U8, I8, F32, F16 = 0, 1, 2, 3
L2, IP = 1, 2
MAX_LEVELS = 16

_NP_DTYPE = {U8: np.uint8, I8: np.int8, F32: np.float32, F16: np.float16}
_NAME = {U8: "u8", I8: "i8", F32: "f32", F16: "f16"}


def dtype_code(arr_dtype) -> int:
    dt = np.dtype(arr_dtype)
    for code, npd in _NP_DTYPE.items():
        if dt == np.dtype(npd):
            return code
    raise ValueError(f"unsupported dtype {dt}")


class Entry(C.Structure):
    _fields_ = [("dist", C.c_float), ("id", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [
        ("hops", C.c_int64 * MAX_LEVELS),
        ("neighbors", C.c_int64 * MAX_LEVELS),
        ("dists", C.c_int64 * MAX_LEVELS),
    ]

    def as_dict(self):
        return {
            "hops": list(self.hops),
            "neighbors": list(self.neighbors),
            "dists": list(self.dists),
        }


_lib = None


def build_lib() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build_lib()
        L = C.CDLL(LIB_PATH)
        vp, i32p, f32p, i64 = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int64
        L.orc_distance.restype = C.c_float
        L.orc_distance.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int]
        L.orc_heap_script.restype = C.c_int
        L.orc_heap_script.argtypes = [C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, C.c_int]
        L.orc_cumulative_probability.restype = C.c_int
        L.orc_cumulative_probability.argtypes = [C.c_int, vp, C.c_int]
        L.orc_sample_levels.restype = None
        L.orc_sample_levels.argtypes = [i64, C.c_int, C.c_uint64, vp]
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i64, vp, vp]
        L.orc_destroy.restype = None
        L.orc_destroy.argtypes = [vp]
        L.orc_insert.restype = C.c_int
        L.orc_insert.argtypes = [vp, C.c_int32]
        L.orc_build.restype = C.c_int
        L.orc_build.argtypes = [vp, vp, i64]
        L.orc_search.restype = C.c_int
        L.orc_search.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.orc_search_batch.restype = C.c_int
        L.orc_search_batch.argtypes = [vp, C.c_int, C.c_int, vp, i64, vp, vp, vp, vp]
        for name in ("orc_nlayers", "orc_enter_level"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp]
        L.orc_layer_capacity.restype = i64
        L.orc_layer_capacity.argtypes = [vp, C.c_int]
        L.orc_layer_max_degree.restype = C.c_int
        L.orc_layer_max_degree.argtypes = [vp, C.c_int]
        L.orc_layer_next_id.restype = C.c_int32
        L.orc_layer_next_id.argtypes = [vp, C.c_int]
        for name in ("orc_layer_adj", "orc_layer_dist", "orc_layer_degree", "orc_layer_heuristic",
                     "orc_layer_ext", "orc_layer_data", "orc_layer_has_data"):
            getattr(L, name).restype = vp
            getattr(L, name).argtypes = [vp, C.c_int]
        L.orc_set_layer.restype = C.c_int
        L.orc_set_layer.argtypes = [vp, C.c_int, i64, C.c_int, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
        L.orc_set_enter_level.restype = C.c_int
        L.orc_set_enter_level.argtypes = [vp, C.c_int]
        L.orc_layer_invalid_count.restype = i64
        L.orc_layer_invalid_count.argtypes = [vp, C.c_int]
        L.orc_add_node.restype = C.c_int
        L.orc_add_node.argtypes = [vp, C.c_int, C.c_int32, vp, C.c_int, vp, C.c_int, vp]
        L.orc_add_edge.restype = C.c_int
        L.orc_add_edge.argtypes = [vp, C.c_int, C.c_int32, C.c_int32, C.c_float]
        L.orc_new_in_id.restype = C.c_int32
        L.orc_new_in_id.argtypes = [vp, C.c_int, C.c_int32]
        L.orc_brute_force.restype = None
        L.orc_brute_force.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64, vp, i64, C.c_int, vp, vp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def distance(a: np.ndarray, b: np.ndarray, metric: int) -> float:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    return float(lib().orc_distance(dtype_code(a.dtype), metric, _ptr(a), _ptr(b), a.shape[-1]))


def heap_script(is_max: bool, ops, dists, ids, final_k: int = -1):
    ops = np.ascontiguousarray(ops, dtype=np.int8)
    dists = np.ascontiguousarray(dists, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    cap = len(ops) + max(final_k, 0) + 1
    out = (Entry * cap)()
    n = lib().orc_heap_script(int(is_max), _ptr(ops), len(ops), _ptr(dists), _ptr(ids), final_k,
                              C.cast(out, C.c_void_p), cap)
    return [(out[i].dist, out[i].id) for i in range(min(n, cap))]


def cumulative_probability(M: int) -> np.ndarray:
    out = np.zeros(64, dtype=np.float64)
    n = lib().orc_cumulative_probability(M, _ptr(out), 64)
    return out[:n].copy()


def sample_levels(n: int, M: int, seed: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint8)
    lib().orc_sample_levels(n, M, seed, _ptr(out))
    return out


def brute_force(base: np.ndarray, queries: np.ndarray, k: int, metric: int):
    base = np.ascontiguousarray(base)
    queries = np.ascontiguousarray(queries, dtype=base.dtype)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), dtype=np.int32)
    ds = np.zeros((nq, k), dtype=np.float64)
    lib().orc_brute_force(dtype_code(base.dtype), metric, base.shape[1], _ptr(base), base.shape[0],
                          _ptr(queries), nq, k, _ptr(ids), _ptr(ds))
    return ids, ds


class Graph:
    """pickle::v2::HNSWGraph restated (oracle.c).  levels are per external id."""

    def __init__(self, base: np.ndarray, metric: int, M: int, efc: int, levels: np.ndarray):
        self.base = np.ascontiguousarray(base)
        self.dtype = dtype_code(self.base.dtype)
        self.metric = metric
        self.dim = self.base.shape[1]
        self.M = M
        self.efc = efc
        self.levels = np.ascontiguousarray(levels, dtype=np.uint8)
        self.h = lib().orc_create(self.dtype, metric, self.dim, M, efc, self.base.shape[0],
                                  _ptr(self.base), _ptr(self.levels))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def build(self, order: np.ndarray) -> None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        if lib().orc_build(self.h, _ptr(order), len(order)) != 0:
            raise RuntimeError("oracle build failed")

    def insert(self, ext_id: int) -> None:
        if lib().orc_insert(self.h, int(ext_id)) != 0:
            raise RuntimeError("oracle insert failed")

    def search(self, queries: np.ndarray, k: int, ef: int, counters: bool = False):
        q = np.ascontiguousarray(queries, dtype=self.base.dtype)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        ids = np.zeros((nq, k), dtype=np.int32)
        ds = np.zeros((nq, k), dtype=np.float32)
        ns = np.zeros(nq, dtype=np.int32)
        ctr = Counters()
        lib().orc_search_batch(self.h, k, ef, _ptr(q), nq, _ptr(ids), _ptr(ds), _ptr(ns),
                               C.byref(ctr) if counters else None)
        if counters:
            return ids, ds, ns, ctr.as_dict()
        return ids, ds, ns

    @property
    def nlayers(self) -> int:
        return lib().orc_nlayers(self.h)

    @property
    def enter_level(self) -> int:
        return lib().orc_enter_level(self.h)

    def layer(self, level: int) -> dict:
        """Copies of a layer's arrays (adj/dist include the slack row)."""
        L = lib()
        cap = L.orc_layer_capacity(self.h, level)
        D = L.orc_layer_max_degree(self.h, level)
        row_bytes = self.base.dtype.itemsize * self.dim

        def arr(fn, dtype, count):
            p = getattr(L, fn)(self.h, level)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                         shape=(count,)).copy()

        return {
            "cap": cap,
            "D": D,
            "next_id": L.orc_layer_next_id(self.h, level),
            "adj": arr("orc_layer_adj", np.int32, (cap + 1) * D).reshape(cap + 1, D),
            "dist": arr("orc_layer_dist", np.float32, (cap + 1) * D).reshape(cap + 1, D),
            "deg": arr("orc_layer_degree", np.int16, cap),
            "heur": arr("orc_layer_heuristic", np.int16, cap),
            "ext": arr("orc_layer_ext", np.int32, cap),
            "data": arr("orc_layer_data", np.uint8, cap * row_bytes).reshape(cap, row_bytes),
            "has_data": arr("orc_layer_has_data", np.uint8, cap),
        }

    def set_layer(self, level: int, layer: dict) -> None:
        cap, D = int(layer["cap"]), int(layer["D"])
        adj = np.ascontiguousarray(layer["adj"][:cap], dtype=np.int32)
        dist = np.ascontiguousarray(layer["dist"][:cap], dtype=np.float32)
        deg = np.ascontiguousarray(layer["deg"], dtype=np.int16)
        heur = np.ascontiguousarray(layer["heur"], dtype=np.int16)
        ext = np.ascontiguousarray(layer["ext"], dtype=np.int32)
        data = np.ascontiguousarray(layer["data"], dtype=np.uint8)
        has = np.ascontiguousarray(layer["has_data"], dtype=np.uint8)
        rc = lib().orc_set_layer(self.h, level, cap, D, int(layer["next_id"]), _ptr(adj), _ptr(dist),
                                 _ptr(deg), _ptr(heur), _ptr(ext), _ptr(data), _ptr(has))
        if rc != 0:
            raise RuntimeError("orc_set_layer failed")

    def set_enter_level(self, level: int) -> None:
        if lib().orc_set_enter_level(self.h, level) != 0:
            raise RuntimeError("bad enter level")

    def invalid_count(self, level: int) -> int:
        return lib().orc_layer_invalid_count(self.h, level)

    # standalone layer mutation
    def new_in_id(self, level: int, ext_id: int) -> int:
        return lib().orc_new_in_id(self.h, level, ext_id)

    def add_node(self, level, vid, R, P, vdata):
        Ra = (Entry * max(len(R), 1))(*[Entry(d, i) for d, i in R])
        Pa = (Entry * max(len(P), 1))(*[Entry(d, i) for d, i in P])
        v = np.ascontiguousarray(vdata, dtype=self.base.dtype)
        lib().orc_add_node(self.h, level, vid, C.cast(Ra, C.c_void_p), len(R), C.cast(Pa, C.c_void_p),
                           len(P), _ptr(v))

    def add_edge(self, level, vid, nid, dist):
        lib().orc_add_edge(self.h, level, vid, nid, float(dist))


def ref_available() -> bool:
    return os.path.exists(REF_BIN)
