"""pyann.HNSWIndex over the MI355X index (pybind/hnsw.cpp:19-127).

Same names, defaults and argument meaning as the reference module:

    idx = HNSWIndex()            # search_ef=100, build_ef=100, top_k=10, M=32
    idx.build(data)              # 2-D uint8 / int8 / float16 / float32, L2
    ids = idx.search(queries)    # int32 [nq, top_k] of row indices
    idx.save(path); idx.load(path)

Differences a caller can observe:
  * the graph is pickle's v2 HNSW (include/hnsw.hpp) built and searched on
    the GPU; the reference module wraps the older NSW-layer builder;
  * ``num_layer`` is the number of graph layers (the reference binds it to
    ``getTopK``, pybind/hnsw.cpp:123);
  * a failed assertion of the reference (ALWAYS_ASSERT, include/marco.hpp:
    20-27) is a ValueError / RuntimeError here;
  * ``load`` needs the base vectors: those of the last ``build`` or the
    ``data`` argument (the file holds the graph only).
"""
from __future__ import annotations

import numpy as np

from . import abi

_DTYPES = (np.uint8, np.int8, np.float16, np.float32)


class HNSWIndex:
    def __init__(self):
        self._search_ef = 100
        self._build_ef = 100
        self._top_k = 10
        self._M = 32
        self._data: np.ndarray | None = None
        self._index: abi.Index | None = None
        self._index_params = None
        self._fast = False

    # properties (pybind/hnsw.cpp:31-45, 118-123)
    @staticmethod
    def _positive(name, v):
        v = int(v)
        if v < 1:
            raise ValueError(f"{name} must be positive")
        return v

    @property
    def search_ef(self) -> int:
        return self._search_ef

    @search_ef.setter
    def search_ef(self, v):
        self._search_ef = self._positive("search_ef", v)

    @property
    def build_ef(self) -> int:
        return self._build_ef

    @build_ef.setter
    def build_ef(self, v):
        self._build_ef = self._positive("build_ef", v)

    @property
    def top_k(self) -> int:
        return self._top_k

    @top_k.setter
    def top_k(self, v):
        self._top_k = self._positive("top_k", v)

    @property
    def M(self) -> int:
        return self._M

    @M.setter
    def M(self, v):
        self._M = self._positive("M", v)

    @property
    def fast_search(self) -> bool:
        """Search with the fast mode (pk_search_fast: recall-equivalent to
        the reference's search, not bit-exact; search_ef <= 512).  Off by
        default; no counterpart in the reference module."""
        return self._fast

    @fast_search.setter
    def fast_search(self, v: bool):
        self._fast = bool(v)

    @property
    def num_layer(self) -> int:
        return 0 if self._index is None else self._index.info()["nlayers"]

    @staticmethod
    def _check_data(data) -> np.ndarray:
        data = np.asarray(data)
        if data.ndim != 2 or data.shape[0] == 0 or data.shape[1] == 0:
            raise ValueError("data must be a non-empty 2-D array")
        if data.dtype.type not in _DTYPES:
            raise ValueError(f"unsupported dtype {data.dtype} (uint8, int8, float16, float32)")
        return np.ascontiguousarray(data)

    def _new_index(self, data: np.ndarray) -> abi.Index:
        if self._index is not None:
            self._index.close()
        self._index = abi.Index(data.dtype, abi.L2, data.shape[1], self._M, self._build_ef, data.shape[0])
        self._index_params = (self._M, self._build_ef)
        return self._index

    # pybind/hnsw.cpp:48-71
    def build(self, data, seed: int | None = None, max_batch: int = 4096) -> None:
        """Builds an L2 graph over the rows of ``data``; the rows are inserted
        in a random order (getRandomIndices, include/util.hpp:66-80) and row i
        is external id i."""
        data = self._check_data(data)
        rng = np.random.default_rng(seed)
        order = rng.permutation(data.shape[0]).astype(np.int32)
        idx = self._new_index(data)
        level_seed = int(rng.integers(0, 2**63 - 1))
        idx.build(data, order=order, max_batch=max_batch, exact_rows=True, seed=level_seed)
        self._data = data

    # pybind/hnsw.cpp:73-101
    def search(self, query) -> np.ndarray:
        """int32 [nq, top_k] row indices of the nearest base rows."""
        if self._index is None or self._data is None:
            raise RuntimeError("search before build / load")
        q = np.asarray(query)
        if q.ndim != 2 or q.shape[0] == 0:
            raise ValueError("query must be a non-empty 2-D array")
        if q.shape[1] != self._data.shape[1]:
            raise ValueError(f"query width {q.shape[1]} != data width {self._data.shape[1]}")
        if q.dtype != self._data.dtype:
            raise ValueError(f"query dtype {q.dtype} != data dtype {self._data.dtype}")
        run = self._index.search_fast if self._fast else self._index.search
        ids, _, cnt = run(np.ascontiguousarray(q), self._top_k, self._search_ef)
        if (cnt != self._top_k).any():
            raise RuntimeError(f"{int((cnt != self._top_k).sum())} queries returned fewer than top_k results")
        return ids

    # pybind/hnsw.cpp:103-109
    def save(self, filename: str) -> None:
        if self._index is None:
            raise RuntimeError("save before build / load")
        self._index.save_file(filename)

    def load(self, filename: str, data=None) -> None:
        if data is not None:
            data = self._check_data(data)
        elif self._data is not None:
            data = self._data
        else:
            raise RuntimeError("load needs the base vectors (build first or pass data=)")
        idx = self._new_index(data)
        idx.set_base(data)
        idx.load_file(filename)
        self._data = data
