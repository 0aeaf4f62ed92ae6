"""Writes the golden fixtures in this directory from the reference itself.

Runs oracle/_ref/pickle_ref (the reference's own headers, compiled by
oracle/Makefile from the reference sources) on small seeded inputs and
records its outputs.  Needs the reference sources to have been built; it is
run by hand in the container that has them:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  dist_kat.npz     distance known answers, 4 dtypes x L2/IP x several dims
  heap_kat.npz     MinHeap / MaxHeap pop order on tie-heavy scripts
  layer_kat.npz    AddNode / AddEdge row states after every step, including
                   the write one past a full row
  graph_<case>.npz reference build (levels it drew, graph digest, layer-0
                   arrays) and AnnSearch results + profiler counters

Base vectors are not stored: they are regenerated from the seeded generator
in pickle_amd.datasets and checked against the stored SHA-256.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pickle_amd import datasets, serializer  # noqa: E402

REF = O.REF_BIN

# name: (dtype, metric, generator, n, nq, dim, M, efc, efs, extra searches)
GRAPH_CASES = {
    "u8_l2": ("u8", "l2", ("gm_u8", 11, 32), 4000, 100, 128, 16, 64, [10, 16, 32, 64], [(1, 10), (100, 128)]),
    "i8_l2": ("i8", "l2", ("gm_i8", 12, 16), 3000, 100, 100, 8, 40, [10, 20, 40], [(1, 10)]),
    "u8_ip": ("u8", "ip", ("gm_u8", 13, 16), 3000, 80, 64, 8, 40, [10, 20, 40], []),
    "i8_ip": ("i8", "ip", ("gm_i8", 14, 16), 3000, 80, 64, 8, 40, [10, 20, 40], []),
    "f32_l2": ("f32", "l2", ("gm_rounded", 15, 32), 3000, 100, 128, 16, 64, [10, 20, 40, 80], [(100, 128)]),
    "f32_ip": ("f32", "ip", ("vmf", 16, 32), 3000, 100, 96, 12, 48, [10, 20, 40], [(1, 16)]),
    "f16_l2": ("f16", "l2", ("gm_f16", 17, 16), 2000, 80, 64, 8, 40, [10, 20, 40], []),
    "f16_ip": ("f16", "ip", ("vmf_f16", 18, 16), 2000, 80, 64, 8, 40, [10, 20, 40], []),
}


def make_vectors(gen, n, nq, dim):
    kind, seed, ncl = gen
    if kind == "gm_u8":
        return datasets.gaussian_mixture(n, nq, dim, ncl, seed, kind="u8")
    if kind == "gm_i8":
        b, q = datasets.gaussian_mixture(n, nq, dim, ncl, seed, kind="u8")
        return (b.astype(np.int16) - 128).astype(np.int8), (q.astype(np.int16) - 128).astype(np.int8)
    if kind == "gm_rounded":
        return datasets.gaussian_mixture(n, nq, dim, ncl, seed, kind="rounded_f32")
    if kind == "gm_f16":
        b, q = datasets.gaussian_mixture(n, nq, dim, ncl, seed, kind="f32")
        return (b / 16).astype(np.float16), (q / 16).astype(np.float16)
    if kind == "vmf":
        return datasets.vmf_mixture(n, nq, dim, ncl, seed)
    if kind == "vmf_f16":
        b, q = datasets.vmf_mixture(n, nq, dim, ncl, seed)
        return b.astype(np.float16), q.astype(np.float16)
    raise ValueError(kind)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def graph_digest(g: serializer.GraphFile) -> str:
    h = hashlib.sha256()
    for L in g.layers:
        for a in (L.ext, L.deg, L.heur, L.adj, L.dist):
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def read_results(path, nq, k):
    raw = np.fromfile(path, dtype=np.uint8).reshape(nq, 4 + 8 * k)
    cnt = raw[:, :4].copy().view(np.int32)[:, 0]
    pairs = raw[:, 4:].copy().reshape(nq, k, 8)
    ids = pairs[:, :, :4].copy().view(np.int32)[:, :, 0]
    ds = pairs[:, :, 4:].copy().view(np.float32)[:, :, 0]
    return ids, ds, cnt


def run(args):
    r = subprocess.run([REF] + [str(a) for a in args], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
    return r.stdout


def graph_case(tmp, name, spec):
    dtype, metric, gen, n, nq, dim, M, efc, efs, extra = spec
    base, queries = make_vectors(gen, n, nq, dim)
    bpath, qpath, gpath = (os.path.join(tmp, f"{name}.{s}") for s in ("base", "q", "graph"))
    base.tofile(bpath)
    queries.tofile(qpath)
    run(["build", "--base", bpath, "--n", n, "--dim", dim, "--dtype", dtype, "--metric", metric,
         "--M", M, "--efc", efc, "--threads", 1, "--out", gpath])
    g = serializer.read(gpath)
    out = dict(
        params=np.array(json.dumps(dict(dtype=dtype, metric=metric, gen=list(gen), n=n, nq=nq, dim=dim,
                                        M=M, efc=efc, order="identity"))),
        base_sha=np.array(sha(base)), queries_sha=np.array(sha(queries)),
        levels=g.levels, enter_level=np.int32(g.enter_level), graph_sha=np.array(graph_digest(g)),
        nlayers=np.int32(len(g.layers)),
        l0_ext=g.layers[0].ext, l0_deg=g.layers[0].deg, l0_heur=g.layers[0].heur,
        l0_adj=g.layers[0].adj, l0_dist=g.layers[0].dist,
    )
    searches = [(10, ef) for ef in efs] + list(extra)
    for k, ef in searches:
        rpath = os.path.join(tmp, f"{name}.res.k{k}")
        stats = json.loads(run(["search", "--graph", gpath, "--base", bpath, "--queries", qpath, "--n", n,
                                "--nq", nq, "--dim", dim, "--dtype", dtype, "--k", k, "--ef", ef,
                                "--out", rpath, "--counters", 1]))[0]
        ids, ds, cnt = read_results(f"{rpath}.ef{ef}", nq, k)
        tag = f"k{k}_ef{ef}"
        out[f"ids_{tag}"] = ids
        out[f"dists_{tag}"] = ds
        out[f"n_{tag}"] = cnt
        out[f"counters_{tag}"] = np.array([stats["hops"], stats["hops0"], stats["dists"], stats["dists0"],
                                           stats["neighbors"]], dtype=np.int64)
    out["searches"] = np.array(searches, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, f"graph_{name}.npz"), **out)
    print(name, "layers", len(g.layers), "max deg l0", int(g.layers[0].deg.max()), "D", g.layers[0].max_degree)


def dist_kat(tmp):
    rng = np.random.default_rng(100)
    out = {}
    for dt in ("u8", "i8", "f16", "f32"):
        for dim in (16, 32, 33, 96, 100, 103, 128, 200):
            n = 64
            if dt == "u8":
                a = rng.integers(0, 256, (n, dim)).astype(np.uint8)
                b = rng.integers(0, 256, (n, dim)).astype(np.uint8)
                a[0] = 255; b[0] = 0  # largest L2
            elif dt == "i8":
                a = rng.integers(-128, 128, (n, dim)).astype(np.int8)
                b = rng.integers(-128, 128, (n, dim)).astype(np.int8)
                a[0] = -128; b[0] = 127
                a[1] = -128; b[1] = -128  # largest IP
            elif dt == "f16":
                a = rng.normal(0, 4, (n, dim)).astype(np.float16)
                b = rng.normal(0, 4, (n, dim)).astype(np.float16)
            else:
                a = rng.normal(0, 1, (n, dim)).astype(np.float32)
                b = rng.normal(0, 1, (n, dim)).astype(np.float32)
            pa, pb, po = (os.path.join(tmp, s) for s in ("a", "b", "o"))
            a.tofile(pa)
            b.tofile(pb)
            for metric in ("l2", "ip"):
                run(["dist", "--dtype", dt, "--metric", metric, "--n", n, "--dim", dim, "--a", pa, "--b", pb,
                     "--out", po])
                key = f"{dt}_{dim}_{metric}"
                out[f"a_{key}"] = a
                out[f"b_{key}"] = b
                out[f"d_{key}"] = np.fromfile(po, np.float32)
    np.savez_compressed(os.path.join(HERE, "dist_kat.npz"), **out)


def heap_kat(tmp):
    rng = np.random.default_rng(101)
    out = {}
    for case in range(12):
        is_max = case % 2
        nops = 400
        # integer-valued distances from a tiny range: ties everywhere
        vals = rng.integers(0, 6 + 4 * (case // 2), nops).astype(np.float32)
        ops = (rng.random(nops) < 0.65).astype(np.int8)
        ops[:20] = 1
        ids = np.arange(nops, dtype=np.int32)
        final_k = int(rng.integers(1, 30))
        script = bytearray()
        script += np.int32(is_max).tobytes() + np.int32(nops + 1).tobytes()
        for i in range(nops):
            script += np.int8(ops[i]).tobytes() + np.float32(vals[i]).tobytes() + np.int32(ids[i]).tobytes()
        script += np.int8(2).tobytes() + np.float32(0).tobytes() + np.int32(final_k).tobytes()
        sp, op = os.path.join(tmp, "hs"), os.path.join(tmp, "ho")
        with open(sp, "wb") as f:
            f.write(bytes(script))
        run(["heap", "--script", sp, "--out", op])
        raw = np.fromfile(op, dtype=np.uint8).reshape(-1, 8)
        out[f"is_max_{case}"] = np.int32(is_max)
        out[f"ops_{case}"] = ops
        out[f"vals_{case}"] = vals
        out[f"final_k_{case}"] = np.int32(final_k)
        out[f"pop_dist_{case}"] = raw[:, :4].copy().view(np.float32)[:, 0]
        out[f"pop_id_{case}"] = raw[:, 4:].copy().view(np.int32)[:, 0]
    np.savez_compressed(os.path.join(HERE, "heap_kat.npz"), **out)


def layer_kat(tmp):
    """AddNode then AddEdge bursts on a small layer, designed so that full
    rows whose entries all survive the heuristic re-run overflow into the
    next row (hnsw_layer.hpp:441-447)."""
    rng = np.random.default_rng(102)
    out = {}
    for case in range(4):
        # near-orthogonal random rows (dim 32) keep every neighbour in the
        # heuristic, so full rows overflow; rows on a line (dim 8) prune hard
        D, cap, dim = 4 + 2 * case, 24, (8 if case % 2 == 0 else 32)
        vecs = rng.normal(0, 1, (cap, dim)).astype(np.float32)
        if case % 2 == 0:
            vecs[:, 1:] *= 0.01
            vecs[:, 0] = np.arange(cap) * (1 if case == 0 else -1) + rng.normal(0, 0.1, cap)
        else:
            # node 0 at the origin, node i on its own axis: every pair of
            # neighbours is farther apart than either is from node 0, so row 0
            # keeps all of them and overflows into row 1
            vecs = np.zeros((cap, dim), np.float32)
            for i in range(1, cap):
                vecs[i, i] = 1.0 + 0.01 * i
        vecs = vecs.astype(np.float32)
        ops = []
        for v in range(cap):
            ops.append((0, v, 0, 0.0, [], []))
        for v in range(cap):
            ops.append((1, v, 0, 0.0, [], []))
        if case % 2 == 1:
            for i in rng.permutation(np.arange(1, cap)):
                d = float(np.sum((vecs[0] - vecs[i]) ** 2, dtype=np.float64))
                ops.append((2, 0, int(i), np.float32(d), [], []))
        for step in range(160):
            v = int(rng.integers(0, cap))
            nb = int(rng.integers(0, cap))
            if nb == v:
                continue
            d = float(np.sum((vecs[v] - vecs[nb]) ** 2, dtype=np.float64))
            ops.append((2, v, nb, np.float32(d), [], []))
        blob = bytearray()
        blob += np.array([D, cap, dim, len(ops)], np.int32).tobytes()
        blob += vecs.tobytes()
        kinds = []
        for kind, a, b, d, R, P in ops:
            blob += np.array([kind, a, b], np.int32).tobytes() + np.float32(d).tobytes()
            blob += np.array([len(R), len(P)], np.int32).tobytes()
            kinds.append((kind, a, b, d))
        sp, op = os.path.join(tmp, "ls"), os.path.join(tmp, "lo")
        with open(sp, "wb") as f:
            f.write(bytes(blob))
        run(["layer", "--script", sp, "--out", op])
        rows = cap + 1
        per = rows * D * 4 * 2 + rows * 2 * 2
        raw = np.fromfile(op, dtype=np.uint8).reshape(len(ops), per)
        last = raw[-1]
        o = 0
        adj = last[o:o + rows * D * 4].copy().view(np.int32).reshape(rows, D); o += rows * D * 4
        dist = last[o:o + rows * D * 4].copy().view(np.float32).reshape(rows, D); o += rows * D * 4
        deg = last[o:o + rows * 2].copy().view(np.int16); o += rows * 2
        heur = last[o:o + rows * 2].copy().view(np.int16)
        degs_all = raw[:, rows * D * 8: rows * D * 8 + rows * 2].copy().view(np.int16)
        out[f"D_{case}"] = np.int32(D)
        out[f"vecs_{case}"] = vecs
        out[f"ops_{case}"] = np.array([(k, a, b) for k, a, b, _ in kinds], np.int32)
        out[f"opd_{case}"] = np.array([d for *_, d in kinds], np.float32)
        out[f"adj_{case}"] = adj
        out[f"dist_{case}"] = dist
        out[f"deg_{case}"] = deg
        out[f"heur_{case}"] = heur
        out[f"degs_all_{case}"] = degs_all
        print("layer case", case, "D", D, "max degree reached", int(degs_all.max()))
    np.savez_compressed(os.path.join(HERE, "layer_kat.npz"), **out)


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/pickle_ref is missing: run `make -C oracle ref` where the reference sources are")
    with tempfile.TemporaryDirectory() as tmp:
        dist_kat(tmp)
        heap_kat(tmp)
        layer_kat(tmp)
        for name, spec in GRAPH_CASES.items():
            graph_case(tmp, name, spec)


if __name__ == "__main__":
    main()
