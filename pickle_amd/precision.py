"""Host-side helpers of the precision form (low-precision traversal copies
with an exact FP32 re-rank; see pk_search_rerank in include/pickle_gpu.h).

The device converts rows and queries itself; these helpers only choose the
per-dimension int8 quantizer from the base vectors and set up the two
indexes (FP32 graph, low-precision traversal copy sharing it)."""
from __future__ import annotations

import numpy as np

from . import abi


def int8_quantizer(base: np.ndarray, metric: int):
    """(offset, scale) per dimension.  L2: the value range is centred and
    spread over [-127, 127]; IP: symmetric (offset 0), so that the int8 inner
    product stays proportional per dimension."""
    x = np.asarray(base, np.float32)
    if metric == abi.L2:
        lo, hi = x.min(axis=0), x.max(axis=0)
        offset = ((lo.astype(np.float64) + hi) / 2).astype(np.float32)
        scale = ((hi.astype(np.float64) - lo) / 254.0).astype(np.float32)
    else:
        offset = np.zeros(x.shape[1], np.float32)
        scale = (np.abs(x).max(axis=0).astype(np.float64) / 127.0).astype(np.float32)
    scale = np.where(scale > 0, scale, np.float32(1.0)).astype(np.float32)
    return offset, scale


def traversal_index(graph: abi.Index, base_f32: np.ndarray, kind: str) -> abi.Index:
    """An index over `graph`'s graph whose traversal copy is FP16 ("f16") or
    per-dimension int8 ("i8") and whose re-rank rows are `base_f32`."""
    dt = np.float16 if kind == "f16" else np.int8
    idx = abi.Index(dt, graph.metric, graph.dim, graph.M, graph.efc, graph.capacity)
    if kind == "i8":
        idx.set_quantizer(*int8_quantizer(base_f32, graph.metric))
    idx.set_full_rows(base_f32)
    idx.copy_graph(graph)
    return idx
