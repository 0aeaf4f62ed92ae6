"""The graph file layout (include/serializer.hpp:27-55) and its .v2meta
sidecar, read and written by pickle_amd/serializer.py, on the CPU.

A graph built by the reference (compiled from its own sources, oracle/_ref)
is written by the reference harness; the Python reader must see a well-formed
graph in it, and writing what it read must give back both files byte for
byte.  Corrupt files raise instead of yielding a graph."""
import os
import struct
import subprocess

import numpy as np
import pytest

import bench
from pickle_amd import serializer


@pytest.fixture(scope="module")
def ref_graph(tmp_path_factory):
    binary, _ = bench.ref_binary()
    if binary is None:
        pytest.skip("the reference harness (oracle/_ref) is not built")
    d = tmp_path_factory.mktemp("ref")
    n, dim, M = 2000, 16, 8
    base = np.random.default_rng(5).integers(0, 256, (n, dim)).astype(np.uint8)
    base.tofile(d / "base.u8")
    out = str(d / "g.bin")
    bench.cpu_build(binary, str(d / "base.u8"), n, dim, "u8", "l2", M, 40, 1, out)
    return out, n, M


def test_reference_file_reads_as_a_graph(ref_graph):
    path, n, M = ref_graph
    g = serializer.read(path)
    assert g.n_total == n and g.M == M and g.dim == 16
    assert 1 <= len(g.layers) <= 16 and 0 <= g.enter_level < len(g.layers)
    for lvl, L in enumerate(g.layers):
        assert L.adj.shape == (L.n, 2 * M if lvl == 0 else M)
        live = L.ext[: L.next_id]
        assert ((live >= 0) & (live < n)).all() and len(np.unique(live)) == len(live)
        assert (L.deg[: L.next_id] >= 0).all()
    assert g.layers[0].next_id == n  # every row is in layer 0
    uppers = [set(L.ext[: L.next_id].tolist()) for L in g.layers]
    assert all(a >= b for a, b in zip(uppers, uppers[1:]))  # each layer's nodes are in the one below


def test_write_gives_back_the_reference_bytes(ref_graph, tmp_path):
    path, _, _ = ref_graph
    g = serializer.read(path)
    out = str(tmp_path / "again.bin")
    serializer.write(out, g)
    assert open(out, "rb").read() == open(path, "rb").read()
    assert open(out + ".v2meta", "rb").read() == open(path + ".v2meta", "rb").read()


def test_corrupt_files_raise(ref_graph, tmp_path):
    path, _, _ = ref_graph
    raw, meta = open(path, "rb").read(), open(path + ".v2meta", "rb").read()

    def read(data, sidecar=meta):
        p = str(tmp_path / "bad.bin")
        open(p, "wb").write(data)
        open(p + ".v2meta", "wb").write(sidecar)
        return serializer.read(p)

    with pytest.raises((ValueError, struct.error)):
        read(raw[: len(raw) // 2])  # cut short
    with pytest.raises(ValueError):
        read(raw[:8] + struct.pack("<3Q", 24, 1 << 62, 1) + raw[32:])  # size field vs shape
    with pytest.raises(ValueError):
        read(struct.pack("<Q", struct.unpack_from("<Q", raw)[0] + 1) + raw[8:])  # layer count vs sidecar
    with pytest.raises(ValueError):
        read(raw, b"XXXXXXXX" + meta[8:])  # sidecar magic
    assert os.path.exists(path)
