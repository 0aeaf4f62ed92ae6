"""The synthetic workloads: the streamed form of a Gaussian-mixture workload
(rows generated while earlier chunks are inserted) yields exactly the rows and
queries of the in-memory form, chunk boundaries included."""
import numpy as np
import pytest

from pickle_amd import datasets


@pytest.mark.parametrize("name,n,nq,rows", [("gm100m_u8", 200_000, 300, 1 << 16), ("gm1m_u8", 70_000, 50, 1 << 17),
                                            ("gm10m_f32", 100_000, 20, 1 << 16)])
def test_stream_equals_workload(name, n, nq, rows):
    base, queries, params = datasets.workload(name, n=n, nq=nq)
    st, params2 = datasets.workload_stream(name, n=n, nq=nq, rows=rows)
    chunks = list(st.base_chunks())
    assert all(c.shape[0] == rows for c in chunks[:-1])
    np.testing.assert_array_equal(np.concatenate(chunks), base)
    np.testing.assert_array_equal(st.queries(), queries)
    assert params == params2


def test_stream_rules():
    st, _ = datasets.workload_stream("gm1m_u8", n=1000, nq=5, rows=1 << 16)
    with pytest.raises(RuntimeError):
        st.queries()  # the queries come after every base row
    with pytest.raises(ValueError):
        datasets.workload_stream("vmf1m_f32")
