"""ctypes binding of the C-ABI in include/pickle_gpu.h (libpickle_amd.so).

There is no CPU fallback: if the library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpickle_amd.so")

UINT8, INT8, FLOAT32, FLOAT16 = 0, 1, 2, 3
L2, IP = 1, 2
MAX_LEVELS = 16

_DTYPES = {UINT8: np.uint8, INT8: np.int8, FLOAT32: np.float32, FLOAT16: np.float16}

STATUS = {0: "PK_OK", -1: "PK_ERR_ARG", -2: "PK_ERR_HIP", -3: "PK_ERR_NOMEM", -4: "PK_ERR_STATE",
          -5: "PK_ERR_IO", -6: "PK_ERR_CAPACITY", -7: "PK_ERR_DEVICE"}

# every entry point the header declares (checked by the tests)
EXPORTS = [
    "pk_last_error", "pk_version", "pk_index_create", "pk_index_destroy", "pk_index_set_base",
    "pk_index_load_layer", "pk_index_set_enter_level", "pk_index_load_file", "pk_index_save_file",
    "pk_index_info", "pk_index_read_layer", "pk_search", "pk_search_device", "pk_search_device_ex",
    "pk_search_device_static", "pk_search_fast_device", "pk_search_fast", "pk_set_fast_tuning", "pk_search_rerank_fast_device",
    "pk_autotune_fast", "pk_set_build_row_limit",
    "pk_device_error", "pk_set_tuning", "pk_last_launch", "pk_sync", "pk_device_alloc", "pk_device_free",
    "pk_copy_to_device", "pk_copy_to_host", "pk_timer_begin", "pk_timer_end", "pk_brute_force", "pk_wall_clock_khz", "pk_sample_levels", "pk_index_init_build",
    "pk_index_insert", "pk_device_copy", "pk_index_set_quantizer", "pk_index_set_full_rows",
    "pk_index_copy_graph", "pk_search_rerank_device", "pk_search_rerank", "pk_search_graph_create",
    "pk_search_graph_launch", "pk_search_graph_destroy", "pk_autotune_search", "pk_server_create",
    "pk_server_create_ex",
    "pk_server_submit", "pk_server_insert", "pk_server_wait", "pk_server_stats", "pk_server_destroy",
]
ERR_STATE, ERR_DEVICE = -4, -7
BUILD_EXACT_ROWS = 1
BUILD_GRAPH = 2  # pk_index_insert flag: batches run as HIP graphs
SERVE_FAST = 1  # pk_server_create_ex flag: launches run the fast mode
BRUTE_K = 32


class PickleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


COUNTER_DTYPE = np.dtype([("hops", np.int32, 2), ("neighbors", np.int32, 2), ("dists", np.int32, 2),
                          ("latency_ticks", np.int32)])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() or python -m pickle_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t
        sig = {
            "pk_last_error": (C.c_char_p, []),
            "pk_version": (C.c_int, []),
            "pk_index_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i64, C.POINTER(vp)]),
            "pk_index_destroy": (C.c_int, [vp]),
            "pk_index_set_base": (C.c_int, [vp, vp, i64, i64]),
            "pk_index_load_layer": (C.c_int, [vp, C.c_int, i64, C.c_int, i64, vp, vp, vp, vp, vp, vp]),
            "pk_index_set_enter_level": (C.c_int, [vp, C.c_int]),
            "pk_index_load_file": (C.c_int, [vp, C.c_char_p]),
            "pk_index_save_file": (C.c_int, [vp, C.c_char_p]),
            "pk_index_info": (C.c_int, [vp, vp, vp, vp, vp]),
            "pk_index_read_layer": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
            "pk_search": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_device": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_device_ex": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_device_static": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_fast": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_fast_device": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
            "pk_set_fast_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "pk_search_rerank_fast_device": (C.c_int, [vp, vp, i64, i32, i32, i32, vp, vp, vp, vp]),
            "pk_autotune_fast": (C.c_int, [vp, vp, i64, i32, i32, vp, vp]),
            "pk_set_build_row_limit": (C.c_int, [vp, C.c_int]),
            "pk_device_error": (C.c_int, [vp, vp]),
            "pk_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "pk_last_launch": (C.c_int, [vp, vp, vp]),
            "pk_sync": (C.c_int, [vp]),
            "pk_device_alloc": (C.c_int, [C.POINTER(vp), sz]),
            "pk_device_free": (C.c_int, [vp]),
            "pk_copy_to_device": (C.c_int, [vp, vp, sz]),
            "pk_copy_to_host": (C.c_int, [vp, vp, sz]),
            "pk_timer_begin": (C.c_int, [vp]),
            "pk_timer_end": (C.c_int, [vp, C.POINTER(C.c_float)]),
            "pk_brute_force": (C.c_int, [vp, vp, i64, vp, vp]),
            "pk_wall_clock_khz": (C.c_int, [C.POINTER(C.c_int)]),
            "pk_sample_levels": (C.c_int, [i64, C.c_int, C.c_uint64, vp]),
            "pk_index_init_build": (C.c_int, [vp, vp]),
            "pk_index_insert": (C.c_int, [vp, vp, i64, C.c_int, C.c_int]),
            "pk_device_copy": (C.c_int, [vp, vp, vp, sz]),
            "pk_index_set_quantizer": (C.c_int, [vp, vp, vp]),
            "pk_index_set_full_rows": (C.c_int, [vp, vp, i64, i64]),
            "pk_index_copy_graph": (C.c_int, [vp, vp]),
            "pk_search_rerank_device": (C.c_int, [vp, vp, i64, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_rerank": (C.c_int, [vp, vp, i64, i32, i32, i32, vp, vp, vp, vp]),
            "pk_search_graph_create": (C.c_int, [vp, vp, i64, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp,
                                                 C.POINTER(vp)]),
            "pk_search_graph_launch": (C.c_int, [vp]),
            "pk_search_graph_destroy": (C.c_int, [vp]),
            "pk_autotune_search": (C.c_int, [vp, vp, i64, i32, i32, vp, C.POINTER(C.c_float)]),
            "pk_server_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
            "pk_server_create_ex": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
            "pk_server_submit": (C.c_int, [vp, vp, i64, i32, i32, vp, vp, vp, C.POINTER(i64)]),
            "pk_server_insert": (C.c_int, [vp, vp, vp, i64, C.c_int, C.POINTER(i64)]),
            "pk_server_wait": (C.c_int, [vp, i64, i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "pk_server_stats": (C.c_int, [vp, vp]),
            "pk_server_destroy": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise PickleError(rc, lib().pk_last_error().decode())


def wall_clock_khz() -> int:
    k = C.c_int()
    check(lib().pk_wall_clock_khz(C.byref(k)))
    return k.value


def sample_levels(n: int, M: int, seed: int) -> np.ndarray:
    """Insert levels with the reference's distribution (include/util.hpp:38-64)."""
    out = np.zeros(n, np.uint8)
    check(lib().pk_sample_levels(n, M, seed, out.ctypes.data_as(C.c_void_p)))
    return out


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    for code, npd in _DTYPES.items():
        if dt == np.dtype(npd):
            return code
    raise ValueError(f"unsupported dtype {dt}")


class DeviceBuffer:
    """Device memory owned by Python (for HBM-resident search inputs)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().pk_device_alloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().pk_copy_to_device(b.ptr, _p(a), a.nbytes))
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        check(lib().pk_copy_to_host(_p(out), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().pk_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """The serving loop (pk_server_*): submit from any thread, wait per ticket.
    ctypes releases the GIL during the calls, so Python threads submit and
    wait concurrently."""

    def __init__(self, index: "Index", slots: int = 4, slot_queries: int = 4096, max_k: int = 100,
                 max_ef: int = 256, fast: bool = False):
        """fast: launches run the fast mode (PK_SERVE_FAST; max_ef <= 512)."""
        self.index = index
        self.max_k = max_k
        s = C.c_void_p()
        check(lib().pk_server_create_ex(index.h, slots, slot_queries, max_k, max_ef, SERVE_FAST if fast else 0,
                                        C.byref(s)))
        self.s = s
        self._keep = {}

    def submit(self, queries: np.ndarray, k: int, ef: int) -> int:
        q = np.ascontiguousarray(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        ds = np.empty((nq, k), np.float32)
        cnt = np.empty(nq, np.int32)
        t = C.c_int64()
        check(lib().pk_server_submit(self.s, _p(q), nq, k, ef, _p(ids), _p(ds), _p(cnt), C.byref(t)))
        self._keep[t.value] = (ids, ds, cnt)
        return t.value

    def insert(self, ext_ids: np.ndarray, rows: np.ndarray, exact_rows: bool = True) -> int:
        e = np.ascontiguousarray(ext_ids, np.int32)
        r = np.ascontiguousarray(rows)
        t = C.c_int64()
        check(lib().pk_server_insert(self.s, _p(e), _p(r), e.shape[0], BUILD_EXACT_ROWS if exact_rows else 0,
                                     C.byref(t)))
        self._keep[t.value] = None
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(ids, dists, count, latency_us, queue_us); ids etc. are None for inserts."""
        lat, que = C.c_double(), C.c_double()
        rc = lib().pk_server_wait(self.s, ticket, timeout_ms, C.byref(lat), C.byref(que))
        if rc != 0:
            err = PickleError(rc, lib().pk_last_error().decode())
            if rc != ERR_STATE:  # the request finished (failed) and its ticket is released
                res = self._keep.pop(ticket, None)
                if rc == ERR_DEVICE and res is not None:
                    err.partial = res  # (ids, dists, count): count -1 marks the queries that failed
            raise err  # PK_ERR_STATE: timeout or another waiter; the ticket stays valid
        res = self._keep.pop(ticket)
        if res is None:
            return None, None, None, lat.value, que.value
        return res[0], res[1], res[2], lat.value, que.value

    def stats(self) -> dict:
        out = np.zeros(5, np.int64)
        check(lib().pk_server_stats(self.s, _p(out)))
        return dict(zip(["launches", "queries", "inserts", "max_concurrent", "waiting"], out.tolist()))

    def close(self) -> None:
        if self.s:
            check(lib().pk_server_destroy(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SearchGraph:
    """A captured search launch (pk_search_graph_*)."""

    def __init__(self, g, index):
        self.g = g
        self.index = index  # keeps the index alive while the graph exists

    def launch(self) -> None:
        check(lib().pk_search_graph_launch(self.g))

    def close(self) -> None:
        if self.g:
            check(lib().pk_search_graph_destroy(self.g))
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """An HBM-resident HNSW index (one pk_index handle)."""

    def __init__(self, dtype, metric: int, dim: int, M: int, efc: int, capacity: int):
        self.dtype = np.dtype(dtype)
        self.metric = metric
        self.dim = dim
        self.M = M
        self.efc = efc
        self.capacity = capacity
        h = C.c_void_p()
        check(lib().pk_index_create(dtype_code(self.dtype), metric, dim, M, efc, capacity, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().pk_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_base(self, rows: np.ndarray, first: int = 0) -> None:
        rows = np.ascontiguousarray(rows, dtype=self.dtype)
        check(lib().pk_index_set_base(self.h, _p(rows), first, rows.shape[0]))

    def load_layer(self, level: int, ext, degree, adj, has_vector, heur=None, dist=None, D: int | None = None) -> None:
        """adj is [rows, D] (rows >= n; rows past n are slots a degree above D
        reads past the last row) or flat with D given."""
        ext = np.ascontiguousarray(ext, np.int32)
        degree = np.ascontiguousarray(degree, np.int32)
        adj = np.ascontiguousarray(adj, np.int32)
        n = ext.shape[0]
        if D is None:
            if adj.ndim != 2:
                raise ValueError("flat adjacency needs D")
            D = adj.shape[1]
        flat = adj.reshape(-1)
        has = np.ascontiguousarray(has_vector, np.uint8)
        heur = None if heur is None else np.ascontiguousarray(heur, np.int16)
        if dist is not None:
            dist = np.ascontiguousarray(np.asarray(dist, np.float32).reshape(-1)[: n * D])
        check(lib().pk_index_load_layer(self.h, level, n, D, flat.shape[0], _p(ext), _p(degree), _p(heur), _p(flat),
                                        _p(dist), _p(has)))

    def set_enter_level(self, level: int) -> None:
        check(lib().pk_index_set_enter_level(self.h, level))

    def load_file(self, path: str) -> None:
        check(lib().pk_index_load_file(self.h, path.encode()))

    def save_file(self, path: str) -> None:
        check(lib().pk_index_save_file(self.h, path.encode()))

    def info(self):
        nl, el = C.c_int(), C.c_int()
        sizes = np.zeros(MAX_LEVELS, np.int64)
        degs = np.zeros(MAX_LEVELS, np.int32)
        check(lib().pk_index_info(self.h, C.byref(nl), C.byref(el), _p(sizes), _p(degs)))
        return {"nlayers": nl.value, "enter_level": el.value, "sizes": sizes[: nl.value].tolist(),
                "degrees": degs[: nl.value].tolist()}

    def read_layer(self, level: int) -> dict:
        inf = self.info()
        n, D = inf["sizes"][level], inf["degrees"][level]
        out = {"ext": np.zeros(n, np.int32), "deg": np.zeros(n, np.int32), "heur": np.zeros(n, np.int16),
               "adj": np.zeros((n, D), np.int32), "dist": np.zeros((n, D), np.float32)}
        check(lib().pk_index_read_layer(self.h, level, _p(out["ext"]), _p(out["deg"]), _p(out["heur"]),
                                        _p(out["adj"]), _p(out["dist"])))
        return out

    def set_tuning(self, ccap: int = 0, vbits: int = 0, blocks_per_cu: int = 0) -> None:
        check(lib().pk_set_tuning(self.h, ccap, vbits, blocks_per_cu))

    def last_launch(self):
        g, l = C.c_int(), C.c_int()
        check(lib().pk_last_launch(self.h, C.byref(g), C.byref(l)))
        return g.value, l.value

    def search(self, queries: np.ndarray, k, ef, k_stride: int | None = None, counters: bool = False):
        """Batched AnnSearch; k and ef are ints or per-query arrays.

        Returns (ids [nq, k_stride], dists [nq, k_stride], counts [nq]) and, with
        counters=True, a structured array of per-query work counters."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        k_arr = None if np.isscalar(k) else np.ascontiguousarray(k, np.int32)
        ef_arr = None if np.isscalar(ef) else np.ascontiguousarray(ef, np.int32)
        if k_stride is None:
            k_stride = int(k) if k_arr is None else int(k_arr.max())
        ids = np.empty((nq, k_stride), np.int32)
        ds = np.empty((nq, k_stride), np.float32)
        cnt = np.empty(nq, np.int32)
        ctr = np.zeros(nq, COUNTER_DTYPE) if counters else None
        check(lib().pk_search(self.h, _p(q), nq, _p(k_arr), _p(ef_arr), 0 if k_arr is not None else int(k),
                              0 if ef_arr is not None else int(ef), k_stride, _p(ids), _p(ds), _p(cnt), _p(ctr)))
        if counters:
            return ids, ds, cnt, ctr
        return ids, ds, cnt

    def search_device(self, d_queries: DeviceBuffer, nq: int, k: int, ef: int, k_stride: int, d_ids: DeviceBuffer,
                      d_dists: DeviceBuffer, d_count: DeviceBuffer | None = None, d_ctr: DeviceBuffer | None = None,
                      d_k: DeviceBuffer | None = None, d_ef: DeviceBuffer | None = None, ef_max: int | None = None):
        """Enqueue a search whose buffers are all device-resident (no wait)."""
        check(lib().pk_search_device_ex(self.h, d_queries.ptr, nq, d_k.ptr if d_k else None, d_ef.ptr if d_ef else None,
                                        k, ef, ef_max if ef_max else ef, k_stride, d_ids.ptr, d_dists.ptr,
                                        d_count.ptr if d_count else None, d_ctr.ptr if d_ctr else None))

    def search_fast_device(self, d_queries: DeviceBuffer, nq: int, k: int, ef: int, k_stride: int,
                           d_ids: DeviceBuffer, d_dists: DeviceBuffer, d_count: DeviceBuffer | None = None,
                           d_ctr: DeviceBuffer | None = None, d_k: DeviceBuffer | None = None,
                           d_ef: DeviceBuffer | None = None, ef_max: int | None = None):
        """Enqueue a fast-mode search (pk_search_fast_device); no wait."""
        check(lib().pk_search_fast_device(self.h, d_queries.ptr, nq, d_k.ptr if d_k else None,
                                          d_ef.ptr if d_ef else None, k, ef, ef_max if ef_max else ef, k_stride,
                                          d_ids.ptr, d_dists.ptr, d_count.ptr if d_count else None,
                                          d_ctr.ptr if d_ctr else None))

    def set_fast_tuning(self, width: int = 0, hash_bits: int = 0, waves_per_group: int = 0,
                        blocks_per_cu: int = 0) -> None:
        check(lib().pk_set_fast_tuning(self.h, width, hash_bits, waves_per_group, blocks_per_cu))

    def autotune_fast(self, d_queries: DeviceBuffer, nq: int, k: int, ef: int) -> dict:
        out = (C.c_int32 * 4)()
        ms = C.c_float()
        check(lib().pk_autotune_fast(self.h, d_queries.ptr, nq, k, ef, out, C.byref(ms)))
        return {"width": out[0], "hash_bits": out[1], "waves_per_group": out[2], "blocks_per_cu": out[3],
                "ms": ms.value}

    def search_rerank_fast_device(self, d_queries: "DeviceBuffer", nq: int, k: int, ef: int, k_stride: int,
                                  d_ids: "DeviceBuffer", d_dists: "DeviceBuffer", d_count: "DeviceBuffer" = None,
                                  d_counters: "DeviceBuffer" = None) -> None:
        check(lib().pk_search_rerank_fast_device(self.h, d_queries.ptr, nq, k, ef, k_stride, d_ids.ptr, d_dists.ptr,
                                                 d_count.ptr if d_count else None,
                                                 d_counters.ptr if d_counters else None))

    def search_rerank_fast(self, queries: np.ndarray, k: int, ef: int, k_stride: int | None = None):
        """pk_search_rerank_fast_device on host FP32 queries: (ids, dists, counts)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        k_stride = k_stride or k
        bufs = [DeviceBuffer.from_array(q), DeviceBuffer(nq * k_stride * 4), DeviceBuffer(nq * k_stride * 4),
                DeviceBuffer(nq * 4)]
        try:
            self.search_rerank_fast_device(bufs[0], nq, k, ef, k_stride, bufs[1], bufs[2], bufs[3])
            self.sync()
            out = (bufs[1].to_array(np.int32, (nq, k_stride)), bufs[2].to_array(np.float32, (nq, k_stride)),
                   bufs[3].to_array(np.int32, (nq,)))
        finally:
            for b in bufs:
                b.free()
        return out

    def search_fast(self, queries: np.ndarray, k, ef, k_stride: int | None = None, counters: bool = False):
        """Fast-mode search of host queries through pk_search_fast:
        (ids, dists, counts[, counters]); k / ef may be ints or per-query arrays."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        k_arr = None if np.isscalar(k) else np.ascontiguousarray(k, np.int32)
        ef_arr = None if np.isscalar(ef) else np.ascontiguousarray(ef, np.int32)
        if k_stride is None:
            k_stride = int(k) if k_arr is None else int(k_arr.max())
        ids = np.empty((nq, k_stride), np.int32)
        ds = np.empty((nq, k_stride), np.float32)
        cnt = np.empty(nq, np.int32)
        ctr = np.zeros(nq, COUNTER_DTYPE) if counters else None
        check(lib().pk_search_fast(self.h, _p(q), nq, _p(k_arr), _p(ef_arr), 0 if k_arr is not None else int(k),
                                   0 if ef_arr is not None else int(ef), k_stride, _p(ids), _p(ds), _p(cnt),
                                   _p(ctr)))
        if counters:
            return ids, ds, cnt, ctr
        return ids, ds, cnt

    def set_row_limit(self, limit: int) -> None:
        """Longest row a reference-rows build writes (0: 511)."""
        check(lib().pk_set_build_row_limit(self.h, limit))

    def init_build(self, levels: np.ndarray) -> None:
        """InitBuffer: levels for every external id of the capacity."""
        lv = np.ascontiguousarray(levels, np.uint8)
        if lv.shape[0] != self.capacity:
            raise ValueError("levels must cover the whole capacity")
        check(lib().pk_index_init_build(self.h, _p(lv)))

    def insert(self, ext_ids, max_batch: int = 4096, exact_rows: bool = True, graph: bool = False) -> None:
        """graph: each batch's kernel phases run as HIP graphs (PK_BUILD_GRAPH)."""
        ids = np.ascontiguousarray(ext_ids, np.int32)
        flags = (BUILD_EXACT_ROWS if exact_rows else 0) | (BUILD_GRAPH if graph else 0)
        check(lib().pk_index_insert(self.h, _p(ids), ids.shape[0], max_batch, flags))

    def build(self, base: np.ndarray, levels: np.ndarray | None = None, order=None, max_batch: int = 4096,
              exact_rows: bool = True, seed: int = 0, progress=None, chunk: int = 1 << 20,
              graph: bool = False) -> None:
        """Build (include/hnsw.hpp:497-507): base rows, levels, inserts.  With
        `progress`, ids are inserted in chunks and progress(done, total) is
        called after each."""
        self.set_base(base)
        if levels is None:
            levels = sample_levels(self.capacity, self.M, seed)
        self.init_build(levels)
        if order is None:
            order = np.arange(base.shape[0], dtype=np.int32)
        if progress is None:
            self.insert(order, max_batch=max_batch, exact_rows=exact_rows, graph=graph)
            return
        for s in range(0, len(order), chunk):
            self.insert(order[s:s + chunk], max_batch=max_batch, exact_rows=exact_rows, graph=graph)
            progress(min(len(order), s + chunk), len(order))

    def brute_force(self, queries: np.ndarray):
        """The BRUTE_K nearest base rows per query (exact int32 for u8/i8; FP64 for floats, returned as FP32)."""
        q = np.ascontiguousarray(queries, dtype=self.dtype)
        nq = q.shape[0]
        ids = np.empty((nq, BRUTE_K), np.int32)
        ds = np.empty((nq, BRUTE_K), np.float32)
        check(lib().pk_brute_force(self.h, _p(q), nq, _p(ids), _p(ds)))
        return ids, ds

    def sync(self) -> None:
        check(lib().pk_sync(self.h))

    def device_error(self) -> int:
        f = C.c_int()
        check(lib().pk_device_error(self.h, C.byref(f)))
        return f.value

    # ---- HIP graph of one search launch, autotuner ----
    def search_graph(self, d_queries: "DeviceBuffer", nq: int, k, ef, k_stride: int, d_ids: "DeviceBuffer",
                     d_dists: "DeviceBuffer", d_count: "DeviceBuffer" = None, d_counters: "DeviceBuffer" = None,
                     ef_max: int | None = None) -> "SearchGraph":
        """k / ef: ints (one for all queries) or DeviceBuffers of int32 per query
        (then ef_max is required)."""
        dk = k.ptr if isinstance(k, DeviceBuffer) else None
        de = ef.ptr if isinstance(ef, DeviceBuffer) else None
        k_all = 0 if dk else int(k)
        ef_all = 0 if de else int(ef)
        em = ef_max if de else ef_all
        g = C.c_void_p()
        check(lib().pk_search_graph_create(self.h, d_queries.ptr, nq, dk, de, k_all, ef_all, em, k_stride, d_ids.ptr,
                                           d_dists.ptr, d_count.ptr if d_count else None,
                                           d_counters.ptr if d_counters else None, C.byref(g)))
        return SearchGraph(g, self)

    def autotune(self, d_queries: "DeviceBuffer", nq: int, k: int, ef: int):
        t = np.zeros(3, np.int32)
        ms = C.c_float()
        check(lib().pk_autotune_search(self.h, d_queries.ptr, nq, k, ef, _p(t), C.byref(ms)))
        return {"ccap": int(t[0]), "vbits": int(t[1]), "blocks_per_cu": int(t[2]), "ms": ms.value}

    # ---- precision form ----
    def set_quantizer(self, offset: np.ndarray, scale: np.ndarray) -> None:
        off = np.ascontiguousarray(offset, np.float32)
        sc = np.ascontiguousarray(scale, np.float32)
        check(lib().pk_index_set_quantizer(self.h, _p(off), _p(sc)))

    def set_full_rows(self, rows: np.ndarray, first: int = 0) -> None:
        rows = np.ascontiguousarray(rows, np.float32)
        check(lib().pk_index_set_full_rows(self.h, _p(rows), first, rows.shape[0]))

    def copy_graph(self, src: "Index") -> None:
        check(lib().pk_index_copy_graph(self.h, src.h))

    def search_rerank(self, queries: np.ndarray, k: int, ef: int, k_stride: int | None = None,
                      counters: bool = False):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ks = k_stride or k
        ids = np.empty((nq, ks), np.int32)
        ds = np.empty((nq, ks), np.float32)
        cnt = np.empty(nq, np.int32)
        ctr = np.zeros(nq, COUNTER_DTYPE) if counters else None
        check(lib().pk_search_rerank(self.h, _p(q), nq, k, ef, ks, _p(ids), _p(ds), _p(cnt),
                                     _p(ctr) if counters else None))
        return (ids, ds, cnt, ctr) if counters else (ids, ds, cnt)

    def search_rerank_device(self, d_queries: "DeviceBuffer", nq: int, k: int, ef: int, k_stride: int,
                             d_ids: "DeviceBuffer", d_dists: "DeviceBuffer", d_count: "DeviceBuffer" = None,
                             d_counters: "DeviceBuffer" = None) -> None:
        check(lib().pk_search_rerank_device(self.h, d_queries.ptr, nq, k, ef, k_stride, d_ids.ptr, d_dists.ptr,
                                            d_count.ptr if d_count else None,
                                            d_counters.ptr if d_counters else None))

    def device_copy(self, dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> None:
        check(lib().pk_device_copy(self.h, dst.ptr, src.ptr, nbytes))

    def timer_begin(self) -> None:
        check(lib().pk_timer_begin(self.h))

    def timer_end(self) -> float:
        ms = C.c_float()
        check(lib().pk_timer_end(self.h, C.byref(ms)))
        return ms.value
