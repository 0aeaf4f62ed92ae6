"""HIP-graph replay of the search launch and the autotuner, through the C-ABI.

Every replay is compared with the eager result of the same inputs, with other
eager searches run between replays; results must not depend on the tuning."""
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


@pytest.fixture(scope="module")
def u8_index():
    fx, p, base, queries = G.graph_case("u8_l2")
    g = G.oracle_graph(fx, p, base)
    idx = index_from_oracle(g)
    yield idx, queries
    idx.close()


def eager(idx, d_q, nq, ks, efs, k_stride):
    d_k, d_e = abi.DeviceBuffer.from_array(ks), abi.DeviceBuffer.from_array(efs)
    di, dd, dc = abi.DeviceBuffer(nq * k_stride * 4), abi.DeviceBuffer(nq * k_stride * 4), abi.DeviceBuffer(nq * 4)
    from pickle_amd.abi import lib, check
    check(lib().pk_search_device_ex(idx.h, d_q.ptr, nq, d_k.ptr, d_e.ptr, 0, 0, int(efs.max()), k_stride, di.ptr,
                                    dd.ptr, dc.ptr, None))
    idx.sync()
    return (di.to_array(np.int32, (nq, k_stride)), dd.to_array(np.float32, (nq, k_stride)),
            dc.to_array(np.int32, (nq,)))


def test_graph_replays_match_eager(u8_index):
    idx, queries = u8_index
    nq = queries.shape[0]
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 10, 32], nq).astype(np.int32)
    efs = rng.integers(32, 96, nq).astype(np.int32)
    d_q = abi.DeviceBuffer.from_array(queries)
    want = eager(idx, d_q, nq, ks, efs, 32)
    d_k, d_e = abi.DeviceBuffer.from_array(ks), abi.DeviceBuffer.from_array(efs)
    di, dd, dc = abi.DeviceBuffer(nq * 32 * 4), abi.DeviceBuffer(nq * 32 * 4), abi.DeviceBuffer(nq * 4)
    graph = idx.search_graph(d_q, nq, d_k, d_e, 32, di, dd, dc, ef_max=96)
    for rep in range(3):
        say("replay", rep)
        di.to_array(np.int32, (nq, 32))  # host round trip between replays
        idx.search(queries, 10, 40)      # other eager work on the same stream and scratch
        graph.launch()
        idx.sync()
        np.testing.assert_array_equal(di.to_array(np.int32, (nq, 32)), want[0], err_msg=f"replay {rep}")
        np.testing.assert_array_equal(dd.to_array(np.float32, (nq, 32)), want[1], err_msg=f"replay {rep}")
        np.testing.assert_array_equal(dc.to_array(np.int32, (nq,)), want[2], err_msg=f"replay {rep}")
    # new query contents in the same buffer
    q2 = queries[::-1].copy()
    want2 = eager(idx, abi.DeviceBuffer.from_array(q2), nq, ks, efs, 32)
    abi.lib().pk_copy_to_device(d_q.ptr, q2.ctypes.data, q2.nbytes)
    graph.launch()
    idx.sync()
    np.testing.assert_array_equal(di.to_array(np.int32, (nq, 32)), want2[0])
    # while the graph lives, calls that would move its pointers are refused
    idx.set_tuning(0, 20, 0)  # a visited table larger than the graph's scratch
    with pytest.raises(abi.PickleError):
        idx.search(queries, 10, 40)
    graph.close()
    idx.search(queries, 10, 40)
    idx.set_tuning(0, 0, 0)


def test_autotune_keeps_results(u8_index):
    idx, queries = u8_index
    nq = queries.shape[0]
    before = idx.search(queries, 10, 64)
    d_q = abi.DeviceBuffer.from_array(queries)
    t = idx.autotune(d_q, nq, 10, 64)
    say("tuning", t)
    assert t["ms"] > 0 and t["ccap"] >= 0 and t["blocks_per_cu"] >= 0
    after = idx.search(queries, 10, 64)
    for a, b in zip(before, after):
        np.testing.assert_array_equal(a, b)
    idx.set_tuning(0, 0, 0)
