"""pk_search_device_static (scratch in a static device array, no allocation)
against pk_search_device_ex: the same kernels, so ids, distances, counts
and work counters must be identical -- on two calls in a row, the second
with fewer queries, so the array's leftover contents from the first call
are exercised.  Calls go through the C-ABI with device buffers.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from pickle_amd import abi
from tests import golden_cases as G
from tests.gpu_util import index_from_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(idx, fn, q, k, ef, ks=None, efs=None):
    nq = q.shape[0]
    d_q = abi.DeviceBuffer.from_array(np.ascontiguousarray(q))
    d_ids = abi.DeviceBuffer(nq * k * 4)
    d_d = abi.DeviceBuffer(nq * k * 4)
    d_c = abi.DeviceBuffer(nq * 4)
    d_ctr = abi.DeviceBuffer(nq * abi.COUNTER_DTYPE.itemsize)
    d_k = abi.DeviceBuffer.from_array(np.ascontiguousarray(ks, np.int32)) if ks is not None else None
    d_e = abi.DeviceBuffer.from_array(np.ascontiguousarray(efs, np.int32)) if efs is not None else None
    ef_max = int(efs.max()) if efs is not None else ef
    rc = fn(idx.h, d_q.ptr, nq, d_k.ptr if d_k else None,
            d_e.ptr if d_e else None, k, ef, ef_max, k, d_ids.ptr, d_d.ptr,
            d_c.ptr, d_ctr.ptr)
    abi.check(rc)
    idx.sync()
    out = (d_ids.to_array(np.int32, (nq, k)), d_d.to_array(np.float32, (nq, k)), d_c.to_array(np.int32, (nq,)),
           d_ctr.to_array(abi.COUNTER_DTYPE, (nq,)))
    for b in (d_q, d_ids, d_d, d_c, d_ctr, d_k, d_e):
        if b is not None:
            b.free()
    return out


def _same(a, b, tag):
    ids_a, d_a, c_a, ctr_a = a
    ids_b, d_b, c_b, ctr_b = b
    np.testing.assert_array_equal(c_a, c_b, err_msg=tag)
    np.testing.assert_array_equal(ids_a, ids_b, err_msg=tag)
    np.testing.assert_array_equal(d_a, d_b, err_msg=tag)
    for f in ("hops", "neighbors", "dists"):
        np.testing.assert_array_equal(ctr_a[f], ctr_b[f], err_msg=f"{tag} {f}")


@pytest.mark.parametrize("name", ["u8_l2", "f32_ip"])
def test_static_scratch_search_equals_device_search(name):
    fx, p, base, queries = G.graph_case(name)
    g = G.oracle_graph(fx, p, base)
    idx = index_from_oracle(g)
    L = abi.lib()
    rng = np.random.default_rng(3)
    # first call: every query, uniform k / ef; second: fewer queries with
    # per-query k and ef (ef up to 300 needs larger visited tables)
    big = np.concatenate([queries] * 8)
    small = queries[: max(1, len(queries) // 3)]
    ks = rng.integers(1, 11, len(small)).astype(np.int32)
    efs = rng.integers(10, 301, len(small)).astype(np.int32)
    efs = np.maximum(efs, ks)
    say("case", name, "nq", len(big), "then", len(small))
    want1 = _run(idx, L.pk_search_device_ex, big, 10, 64)
    got1 = _run(idx, L.pk_search_device_static, big, 10, 64)
    _same(want1, got1, f"{name} call 1")
    assert (got1[2] == 10).all()
    want2 = _run(idx, L.pk_search_device_ex, small, 10, 0, ks, efs)
    got2 = _run(idx, L.pk_search_device_static, small, 10, 0, ks, efs)
    _same(want2, got2, f"{name} call 2")
    # and the first shape again, after the array held the second call's state
    got3 = _run(idx, L.pk_search_device_static, big, 10, 64)
    _same(want1, got3, f"{name} call 3")
    idx.close()


def test_static_scratch_search_argument_errors():
    fx, p, base, queries = G.graph_case("u8_l2")
    g = G.oracle_graph(fx, p, base)
    idx = index_from_oracle(g)
    L = abi.lib()
    d_q = abi.DeviceBuffer.from_array(np.ascontiguousarray(queries))
    d_o = abi.DeviceBuffer(len(queries) * 10 * 4)
    # ef above the LDS result heap's limit, k above k_stride
    rc = L.pk_search_device_static(idx.h, d_q.ptr, len(queries), None, None, 10, 700, 700, 10,
                                   d_o.ptr, d_o.ptr, None, None)
    assert rc == -1, abi.lib().pk_last_error()
    rc = L.pk_search_device_static(idx.h, d_q.ptr, len(queries), None, None, 20, 64, 64, 10,
                                   d_o.ptr, d_o.ptr, None, None)
    assert rc == -1, abi.lib().pk_last_error()
    # no queries: nothing to do
    rc = L.pk_search_device_static(idx.h, d_q.ptr, 0, None, None, 10, 64, 64, 10,
                                   d_o.ptr, d_o.ptr, None, None)
    assert rc == 0
    d_q.free()
    d_o.free()
    idx.close()
