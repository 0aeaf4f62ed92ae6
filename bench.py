#!/usr/bin/env python3
"""QPS at recall@10 >= 0.95 on one MI355X, beside the reference's CPU search.

Workload (default): 1,000,000 x 128 uint8 vectors from the 64-cluster Gaussian
mixture quantised to [0,255] (seed 1), L2, M=32, efConstruction=256, 10,000
queries, k=10 -- a graph built on the CPU by the reference and searched on
the GPU.

Steps
  1. generate the vectors (pickle_amd.datasets, seeded);
  2. CPU baseline leg: the reference's own HNSWGraph::Build (oracle/_ref, the
     reference headers compiled from source) on all host threads, timed, and
     written in the reference serializer layout;
  3. exhaustive ground truth on the GPU (pk_brute_force, exact int32);
  4. load the graph into HBM, sweep ef, pick the smallest ef whose recall@10
     (label matches, benchmark/hnsw_pickle.cpp:94-112) reaches the target;
  5. time --steps launches of the whole query batch at that ef with HIP events
     on the index stream (inputs and outputs resident in HBM);
  6. CPU baseline leg: the reference's AnnSearch over the same queries, graph,
     ef and k on the host cores.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10>=0.95 (1 GPU vs host CPU); p99 latency; achieved HBM GB/s"
HBM_PEAK_GBS = 8000.0
CALIB_BYTES = 256 << 20
PROFILES = os.path.join(ROOT, "profiles")

WORKLOADS = {
    "gm1m_u8": dict(desc="1M x 128 uint8, 64-cluster gaussian mixture quantised to [0,255] (seed 1), L2, "
                         "M=32, efC=256, 10k queries, k=10, CPU-built graph", n=1_000_000, nq=10_000),
    "gm100k_f32": dict(desc="100k x 128 float32 (rounded), 64-cluster gaussian mixture (seed 0), L2, M=16, "
                            "efC=200, 1k queries, k=10", n=100_000, nq=1_000),
    "gm10m_f32": dict(desc="10M x 128 float32, 64-cluster gaussian mixture (seed 3), L2, M=32, efC=200, 100k "
                           "queries, k=10, GPU-built graph, fp16 traversal copy + FP32 re-rank",
                      n=10_000_000, nq=100_000, build="gpu", precision="f16"),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    return min(os.cpu_count() or 1, 16)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def has_avx512() -> bool:
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd"))


# --------------------------------------------------------------------------
# cpu_baseline leg: the reference compiled from its sources (oracle/_ref)
# --------------------------------------------------------------------------

def ref_binary():
    from oracle import oracle as O

    if has_avx512() and os.path.exists(O.REF_BIN):
        return O.REF_BIN, "avx512"
    if os.path.exists(O.REF_BIN_GENERIC):
        return O.REF_BIN_GENERIC, "generic"
    return None, None


def cpu_build(binary, base_path, n, dim, dtype, metric, M, efc, threads, out):
    cmd = [binary, "build", "--base", base_path, "--n", str(n), "--dim", str(dim), "--dtype", dtype,
           "--metric", metric, "--M", str(M), "--efc", str(efc), "--threads", str(threads), "--out", out,
           "--shuffle", "1"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference build failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_search(binary, graph, base_path, q_path, n, nq, dim, dtype, k, ef, threads, repeat, out):
    cmd = [binary, "search", "--graph", graph, "--base", base_path, "--queries", q_path, "--n", str(n),
           "--nq", str(nq), "--dim", str(dim), "--dtype", dtype, "--k", str(k), "--ef", str(ef),
           "--threads", str(threads), "--repeat", str(repeat), "--out", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference search failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout)[0]


def read_ref_results(path, nq, k):
    raw = np.fromfile(path, dtype=np.uint8).reshape(nq, 4 + 8 * k)
    pairs = raw[:, 4:].copy().reshape(nq, k, 8)
    return pairs[:, :, :4].copy().view(np.int32)[:, :, 0]


# --------------------------------------------------------------------------
# rocprofv3 summaries (profiles/) and the traffic they record
# --------------------------------------------------------------------------

def pmc_profile_path(workload, ef):
    return os.path.join(PROFILES, f"k_search_{workload}_ef{ef}_pmc.md")


def profiled_traffic(workload, n, nq, k, ef):
    """HBM bytes per k_search launch recorded by a --summarize-pmc run of the
    same workload, size and ef (None if there is none)."""
    path = pmc_profile_path(workload, ef)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic-json: "):
            rec = json.loads(line[len("traffic-json: "):])
            if (rec["n"], rec["nq"], rec["k"], rec["ef"]) == (n, nq, k, ef):
                return rec["traffic_bytes"], os.path.relpath(path, ROOT)
    return None, None


def _csv_rows(directory, suffix):
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    rows = []
    for f in files:
        rows.extend(csv.DictReader(open(f)))
    return rows


def _short_kernel(name):
    for key in ("k_search", "k_brute", "k_copy16", "k_gather_rows", "k_fill_i32", "k_insert_search",
                "k_insert_rows", "k_insert_edges", "k_rerank"):
        if key in name:
            return key + (name[name.index("<"):name.index(">") + 1] if "<" in name and key != "k_copy16" else "")
    return name[:60]


def _counter_per_dispatch(directory, counter):
    out = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        out.setdefault(_short_kernel(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return out


def summarize(args):
    """Writes a profiles/ summary from rocprofv3 output and the bench line of
    the same command."""
    bl = json.loads(open(args.bench_json).read().strip().splitlines()[-1])
    cfg, roof = bl["config"], bl["roofline"]
    head = [f"# k_search — {cfg['workload']}", "",
            f"ef={cfg['ef']}, k={cfg['k']}, {cfg['nq']} queries per launch, recall@10 = {cfg['recall@10']}", "",
            "Command (GPU box, committed files only):", ""]
    head += ["    " + c for c in args.cmd]
    body = []
    if args.summarize_trace:
        rows = _csv_rows(args.summarize_trace, "kernel_stats.csv")
        body += ["", "`kernel_stats.csv` (per-kernel totals over the whole process):", "",
                 "| Kernel | Calls | Average | Min | Max |", "|---|---|---|---|---|"]
        search_avg = None
        for r in rows:
            avg = float(r["AverageNs"]) / 1e6
            name = _short_kernel(r["Name"])
            if name.startswith("k_search"):
                search_avg = avg
            body.append(f"| `{name}` | {r['Calls']} | {avg:.4f} ms | {float(r['MinNs']) / 1e6:.4f} ms | "
                        f"{float(r['MaxNs']) / 1e6:.4f} ms |")
        body += ["", f"bench.py in the same run: {roof['kernel_ms']} ms per k_search launch (HIP events on the "
                 f"index stream; trace average {search_avg:.4f} ms), {bl['value']:,.0f} QPS, algorithmic "
                 f"{roof['algorithmic_bytes_per_launch'] / 1e9:.3f} GB per launch -> {roof['achieved']} GB/s "
                 f"= {roof['frac'] * 100:.2f} % of 8 TB/s.",
                 f"Per query: {roof['per_query']['hops']} hops, {roof['per_query']['neighbors']} adjacency "
                 f"entries scanned, {roof['per_query']['dists']} distances."]
    if args.summarize_pmc:
        fdir, wdir = args.summarize_pmc
        fetch = _counter_per_dispatch(fdir, "FETCH_SIZE")
        write = _counter_per_dispatch(wdir, "WRITE_SIZE")
        kib = CALIB_BYTES / 1024.0
        ff = (sum(fetch["k_copy16"]) / len(fetch["k_copy16"])) / kib
        fw = (sum(write["k_copy16"]) / len(write["k_copy16"])) / kib
        skey = [key for key in fetch if key.startswith("k_search")][0]
        f_kib = sum(fetch[skey]) / len(fetch[skey])
        w_kib = sum(write[skey]) / len(write[skey])
        fb, wb = f_kib * 1024 / ff, w_kib * 1024 / fw
        alg = roof["algorithmic_bytes_per_launch"]
        body += ["", "Hardware byte counters (two `--pmc` passes, KiB per dispatch, averaged over the "
                 "dispatches of each kernel):", "",
                 "| Kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB |", "|---|---|---|---|"]
        for key in sorted(set(fetch) | set(write)):
            fv, wv = fetch.get(key, [0]), write.get(key, [0])
            body.append(f"| `{key}` | {len(fv)} | {sum(fv) / len(fv):,.1f} | {sum(wv) / len(wv):,.1f} |")
        body += ["", f"Calibration: `k_copy16` reads and writes {CALIB_BYTES >> 20} MiB "
                 f"({kib:,.0f} KiB); FETCH_SIZE reads {ff:.4f} x and WRITE_SIZE {fw:.4f} x that.",
                 f"k_search per launch: {fb / 1e9:.3f} GB fetched + {wb / 1e9:.3f} GB written = "
                 f"{(fb + wb) / 1e9:.3f} GB of HBM traffic, against {alg / 1e9:.3f} GB algorithmic "
                 f"({(fb + wb) / alg:.2f} x).", ""]
        rec = {"workload": args.workload, "n": cfg["n"], "nq": cfg["nq"], "k": cfg["k"], "ef": cfg["ef"],
               "fetch_bytes": int(fb), "write_bytes": int(wb), "traffic_bytes": int(fb + wb),
               "algorithmic_bytes": alg, "factor_fetch": round(ff, 5), "factor_write": round(fw, 5)}
        body.append("traffic-json: " + json.dumps(rec))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(head + body) + "\n")
    print(f"wrote {args.out}")


def recall_at_k(ids, gt, k):
    """benchmark/hnsw_pickle.cpp:94-112: a returned id counts if it is one of
    the first k ground-truth labels (ties are not credited)."""
    hit = 0
    for i in range(ids.shape[0]):
        hit += len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist()))
    return hit / (k * ids.shape[0])


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gm1m_u8", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="override the base size (smaller runs)")
    ap.add_argument("--nq", type=int, default=None)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--recall-target", type=float, default=0.95)
    ap.add_argument("--efs", default="10,12,14,16,20,24,28,32,40,48,64,96,128,192,256,384,512")
    ap.add_argument("--ef", type=int, default=0, help="fixed ef (skips the sweep)")
    ap.add_argument("--graph-cache", default=None, help="directory to reuse / keep the CPU-built graph")
    ap.add_argument("--no-cpu-search", action="store_true", help="skip the CPU search baseline")
    ap.add_argument("--cpu-repeat", type=int, default=3)
    ap.add_argument("--cpu-queries", type=int, default=0, help="bound the CPU search sample (0 = all queries)")
    ap.add_argument("--gt-queries", type=int, default=0, help="queries with ground truth (0 = all)")
    ap.add_argument("--build", choices=["cpu", "gpu"], default=None,
                    help="graph from the reference's CPU build or the GPU batched build (default per workload)")
    ap.add_argument("--insert-batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", choices=["none", "f16", "i8"], default=None,
                    help="search a low-precision traversal copy + FP32 re-rank (float workloads)")
    ap.add_argument("--summarize-trace", default=None, metavar="DIR",
                    help="write a profiles/ summary from a rocprofv3 --kernel-trace --stats output directory")
    ap.add_argument("--summarize-pmc", nargs=2, default=None, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write a profiles/ summary from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE directories")
    ap.add_argument("--bench-json", default=None, help="the JSON line bench.py printed in the profiled run")
    ap.add_argument("--out", default=None, help="summary file to write")
    ap.add_argument("--cmd", action="append", default=[], help="command lines that produced the profile")
    args = ap.parse_args()

    if args.summarize_trace or args.summarize_pmc:
        summarize(args)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (run with --gpus 1). "
              "Multi-GPU sharding is described in DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    from pickle_amd import abi, datasets

    wl = WORKLOADS[args.workload]
    n = args.n or wl["n"]
    nq = args.nq or wl["nq"]
    k = args.k
    t0 = time.time()
    base, queries, params = datasets.workload(args.workload, n=n, nq=nq)
    dtype_name = {np.dtype(np.uint8): "u8", np.dtype(np.int8): "i8", np.dtype(np.float32): "f32",
                  np.dtype(np.float16): "f16"}[base.dtype]
    dim = base.shape[1]
    metric = params["metric"]
    M, efc = params["M"], params["efc"]
    log(f"workload {args.workload}: n={n} nq={nq} dim={dim} {dtype_name} {metric} M={M} efC={efc} "
        f"(generated in {time.time() - t0:.1f}s)")

    threads = host_threads()
    binary, isa = ref_binary()
    if binary is None:
        sys.exit("the reference build (oracle/_ref) is missing: run __graft_entry__.build() where the "
                 "reference sources are")
    build = args.build or wl.get("build", "cpu")
    prec = args.precision or wl.get("precision", "none")
    if prec != "none" and base.dtype != np.float32:
        sys.exit("--precision needs a float32 workload")

    work = args.graph_cache or tempfile.mkdtemp(prefix="pk_bench_")
    os.makedirs(work, exist_ok=True)
    tag = f"{args.workload}_n{n}_{build}"
    base_path = os.path.join(work, f"{tag}.base")
    q_path = os.path.join(work, f"{tag}_q{nq}.queries")
    graph_path = os.path.join(work, f"{tag}.graph")
    meta_path = graph_path + ".build.json"
    base.tofile(base_path)
    queries.tofile(q_path)
    mcode = abi.L2 if metric == "l2" else abi.IP
    idx = abi.Index(base.dtype, mcode, dim, M, efc, n)
    idx.set_base(base)
    if os.path.exists(graph_path) and os.path.exists(meta_path):
        build_info = json.load(open(meta_path))
        log(f"reusing graph {graph_path} (built in {build_info['build_seconds']:.1f}s)")
        idx.load_file(graph_path)
    elif build == "cpu":
        log(f"reference CPU build ({isa}, {threads} threads) ...")
        build_info = cpu_build(binary, base_path, n, dim, dtype_name, metric, M, efc, threads, graph_path)
        json.dump(build_info, open(meta_path, "w"))
        log(f"built in {build_info['build_seconds']:.1f}s, {build_info['layers']} layers")
        idx.load_file(graph_path)
    else:
        log(f"GPU batched build (insert batches up to {args.insert_batch}) ...")
        tb = time.perf_counter()
        idx.build(base, max_batch=args.insert_batch, seed=args.seed)
        idx.sync()
        build_info = {"build_seconds": time.perf_counter() - tb, "threads": 0, "where": "gpu",
                      "insert_batch": args.insert_batch}
        log(f"GPU build in {build_info['build_seconds']:.1f}s")
        idx.save_file(graph_path)  # for the reference's search on the same graph
        json.dump(build_info, open(meta_path, "w"))
    info = idx.info()
    log(f"graph in HBM: layers {info['sizes']}")

    gt_nq = min(nq, args.gt_queries) if args.gt_queries else nq
    t1 = time.time()
    gt_ids, gt_d = idx.brute_force(queries[:gt_nq])
    if base.dtype.kind == "f":  # FP64 re-rank of the FP32 candidates
        A, Q = base.astype(np.float64), queries[:gt_nq].astype(np.float64)
        for i in range(gt_nq):
            c = gt_ids[i]
            d64 = ((A[c] - Q[i]) ** 2).sum(1) if metric == "l2" else -(A[c] @ Q[i])
            gt_ids[i] = c[np.lexsort((c, d64))]
        del A
    log(f"ground truth (GPU brute force, {gt_nq} queries) in {time.time() - t1:.1f}s")

    # HBM-resident buffers for the timed runs
    d_q = abi.DeviceBuffer.from_array(queries)
    d_ids = abi.DeviceBuffer(nq * k * 4)
    d_dist = abi.DeviceBuffer(nq * k * 4)
    d_cnt = abi.DeviceBuffer(nq * 4)
    d_ctr = abi.DeviceBuffer(nq * abi.COUNTER_DTYPE.itemsize)
    tidx = None
    if prec != "none":
        from pickle_amd import precision

        tp = time.perf_counter()
        tidx = precision.traversal_index(idx, base, prec)
        log(f"{prec} traversal copy + FP32 rows in HBM ({time.perf_counter() - tp:.1f}s)")

    def backend(which):
        if which == "fp32":
            return idx, (lambda ef, c=False: idx.search_device(d_q, nq, k, ef, k, d_ids, d_dist, d_cnt,
                                                               d_ctr if c else None))
        return tidx, (lambda ef, c=False: tidx.search_rerank_device(d_q, nq, k, ef, k, d_ids, d_dist, d_cnt,
                                                                    d_ctr if c else None))

    def sweep(which, target, efs):
        h, run = backend(which)
        out, chosen = [], None
        for ef in efs:
            if ef < k:
                continue
            run(ef)
            h.sync()
            h.timer_begin()
            run(ef)
            ms = h.timer_end()
            ids = d_ids.to_array(np.int32, (nq, k))[:gt_nq]
            r = recall_at_k(ids, gt_ids, k)
            out.append([ef, round(r, 4), round(nq / (ms / 1e3), 1)])
            log(f"[{which}] ef={ef:4d} recall@{k}={r:.4f} qps={nq / (ms / 1e3):,.0f}")
            if r >= target and chosen is None:
                chosen = ef
                if not args.ef:
                    break
        if chosen is None:
            chosen = out[-1][0]
            log(f"[{which}] no ef reached recall {target}; using ef={chosen}")
        return chosen, [x[1] for x in out if x[0] == chosen][0], out

    def timed(which, ef):
        h, run = backend(which)
        for _ in range(args.warmup):
            run(ef)
        h.sync()
        tw0 = time.perf_counter()
        h.timer_begin()
        for _ in range(args.steps):
            run(ef)
        ev_ms = h.timer_end()
        h.sync()
        return time.perf_counter() - tw0, ev_ms / 1e3 / args.steps

    efs = [args.ef] if args.ef else [int(e) for e in args.efs.split(",")]
    comparison = None
    if prec == "none":
        main_be = "fp32"
        chosen, recall, sweep_out = sweep("fp32", args.recall_target, efs)
    else:
        ef32, r32, sw32 = sweep("fp32", args.recall_target, efs)
        wall32, ks32 = timed("fp32", ef32)
        qps32 = nq * args.steps / wall32
        # the precision form at equal recall: the first ef within 0.002 of the FP32 recall
        main_be = "rerank"
        chosen, recall, sweep_out = sweep("rerank", r32 - 0.002, efs)
        comparison = {"fp32_traversal": {"ef": ef32, "recall@10": r32, "qps": round(qps32, 1),
                                         "kernel_ms": round(ks32 * 1e3, 4), "sweep_ef_recall_qps": sw32}}
    h_main, run_main = backend(main_be)
    if h_main.device_error():
        sys.exit("device reported a search failure")

    # work counters (untimed pass, as the reference's second profiler pass)
    run_main(chosen, True)
    h_main.sync()
    ctr = d_ctr.to_array(abi.COUNTER_DTYPE, (nq,))
    elem = {"none": base.dtype.itemsize, "f16": 2, "i8": 1}[prec]
    hops = ctr["hops"].sum(axis=1).astype(np.int64)
    nbr = ctr["neighbors"].sum(axis=1).astype(np.int64)
    dists = ctr["dists"].sum(axis=1).astype(np.int64)
    # algorithmic bytes (reference Profiler units): 4 B per adjacency entry
    # scanned, 8 B of node metadata per hop, one row per distance, the query;
    # the precision form adds the FP32 rows of the ef re-ranked candidates
    alg_bytes = int((4 * nbr + 8 * hops + elem * dim * dists).sum() + nq * elem * dim)
    if prec != "none":
        alg_bytes += nq * (chosen * (4 * dim + 4) + 4 * dim)
    khz = abi.wall_clock_khz()
    lat_ms = ctr["latency_ticks"].astype(np.float64) / khz

    wall, kernel_s = timed(main_be, chosen)
    qps = nq * args.steps / wall
    achieved = alg_bytes / kernel_s / 1e9
    grid, lds = h_main.last_launch()
    log(f"timed: {args.steps} x {nq} queries at ef={chosen}: {qps:,.0f} QPS, kernel {kernel_s * 1e3:.3f} ms, "
        f"{achieved:.1f} GB/s algorithmic (grid {grid}, LDS {lds} B/workgroup)")
    if comparison:
        comparison["speedup_vs_fp32_traversal"] = round(qps / comparison["fp32_traversal"]["qps"], 3)
        log(f"{prec} traversal + FP32 re-rank: {qps:,.0f} QPS at recall {recall:.4f} vs FP32 traversal "
            f"{comparison['fp32_traversal']['qps']:,.0f} QPS at {comparison['fp32_traversal']['recall@10']:.4f}"
            f" -> {comparison['speedup_vs_fp32_traversal']:.3f}x")

    # known-size copy (k_copy16, 256 MiB read + 256 MiB written, twice):
    # the kernel against which profiled HBM byte counters are calibrated
    calib = abi.DeviceBuffer(CALIB_BYTES)
    calib2 = abi.DeviceBuffer(CALIB_BYTES)
    for _ in range(2):
        idx.device_copy(calib2, calib, CALIB_BYTES)
    idx.sync()
    calib.free()
    calib2.free()

    # cpu_baseline leg: the reference's AnnSearch on the host cores, same
    # graph, FP32 data, the FP32 operating point
    cpu = None
    if not args.no_cpu_search:
        cpu_ef = comparison["fp32_traversal"]["ef"] if comparison else chosen
        cnq = min(nq, args.cpu_queries) if args.cpu_queries else nq
        if cnq < nq:
            queries[:cnq].tofile(q_path + ".sample")
            cq_path = q_path + ".sample"
        else:
            cq_path = q_path
        out = os.path.join(work, "cpu_res")
        st = cpu_search(binary, graph_path, base_path, cq_path, n, cnq, dim, dtype_name, k, cpu_ef, threads,
                        args.cpu_repeat, out)
        cpu_ids = read_ref_results(f"{out}.ef{cpu_ef}", cnq, k)
        r_n = min(cnq, gt_nq)
        cpu_recall = recall_at_k(cpu_ids[:r_n], gt_ids[:r_n], k)
        entry = {"value": round(st["qps"], 1), "unit": "queries/s", "cores": threads, "kind": "reference",
                 "sample": f"{cnq} queries, ef={cpu_ef}, k={k}, best of {args.cpu_repeat} runs, reference "
                           f"AnnSearch compiled from source ({isa} distance path) on {cpu_model()}, "
                           f"{'reference-built' if build == 'cpu' else 'GPU-built'} graph",
                 "qps_per_thread": round(st["qps"] / threads, 1), "recall@10": round(cpu_recall, 4),
                 "build_seconds": round(build_info["build_seconds"], 2),
                 "build_threads": build_info["threads"], "build_where": build_info.get("where", "cpu")}
        if not comparison:
            same = float(np.mean(np.all(cpu_ids == d_ids.to_array(np.int32, (nq, k))[:cnq], axis=1)))
            entry["ids_identical_to_gpu"] = round(same, 4)
            log(f"CPU reference: {st['qps']:,.0f} QPS on {threads} threads, recall {cpu_recall:.4f}, "
                f"GPU ids identical on {same * 100:.2f}% of queries")
        else:
            log(f"CPU reference: {st['qps']:,.0f} QPS on {threads} threads at ef={cpu_ef}, recall {cpu_recall:.4f}")
        cpu = entry

    traffic, traffic_src = profiled_traffic(args.workload, n, nq, k, chosen) if prec == "none" else (None, None)
    search_desc = "exact (reference heap order)" if prec == "none" else \
        f"{prec} traversal copy + exact FP32 re-rank of the ef candidates"
    dtype_out = dtype_name if prec == "none" else f"{prec} traversal, f32 re-rank"
    cfg = {"workload": wl["desc"] if (n, nq) == (wl["n"], wl["nq"]) else f"{wl['desc']} (n={n}, nq={nq})",
           "n": n, "nq": nq, "dim": dim, "M": M, "efC": efc, "k": k, "ef": chosen, "recall@10": recall,
           "recall_queries": gt_nq, "graph": "reference CPU build" if build == "cpu" else
           f"GPU batched build ({build_info['build_seconds']:.1f}s)", "search": search_desc,
           "p50_query_latency_ms": round(float(np.percentile(lat_ms, 50)), 4),
           "p99_query_latency_ms": round(float(np.percentile(lat_ms, 99)), 4),
           "batch_latency_ms": round(wall / args.steps * 1e3, 4), "sweep_ef_recall_qps": sweep_out}
    if comparison:
        cfg["precision_comparison"] = comparison
    line = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype_out, "data": "synthetic", "config": cfg,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_search" if prec == "none" else "k_quantize + k_search + k_rerank",
                     "kernel_ms": round(kernel_s * 1e3, 4), "algorithmic_bytes_per_launch": alg_bytes,
                     "per_query": {"hops": round(float(hops.mean()), 2), "neighbors": round(float(nbr.mean()), 2),
                                   "dists": round(float(dists.mean()), 2)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
