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
STREAM_ROWS = 20_000_000  # serving builds of at least this many rows generate them while inserting
PROFILES = os.path.join(ROOT, "profiles")

WORKLOADS = {
    "gm1m_u8": dict(desc="1M x 128 uint8, 64-cluster gaussian mixture quantised to [0,255] (seed 1), L2, "
                         "M=32, efC=256, 10k queries, k=10, CPU-built graph", n=1_000_000, nq=10_000),
    "gm100k_f32": dict(desc="100k x 128 float32 (rounded), 64-cluster gaussian mixture (seed 0), L2, M=16, "
                            "efC=200, 1k queries, k=10", n=100_000, nq=1_000),
    "vmf1m_f32": dict(desc="1M x 96 float32 unit vectors, 256-cluster vMF-like mixture (seed 2), IP, M=24, "
                           "efC=200, 50k queries", n=1_000_000, nq=50_000, build="gpu"),
    "gm100m_u8": dict(desc="100M x 128 uint8, 1024-cluster gaussian mixture quantised to [0,255] (seed 4), L2, "
                           "M=32, efC=256", n=100_000_000, nq=100_000, build="gpu"),
    "gm10m_f32": dict(desc="10M x 128 float32, 64-cluster gaussian mixture (seed 3), L2, M=32, efC=200, 100k "
                           "queries, k=10, GPU-built graph (reference rows up to 128 ids); FP32 exact and fast legs, "
                           "FP16 and per-dimension int8 traversal copies + FP32 re-rank",
                      n=10_000_000, nq=100_000, build="gpu", precision="f16,i8",
                      efs="32,64,96,128,160,192,224,256,288,320,352,384,448,512", gt_queries=10_000,
                      cpu_queries=10_000, row_limit=128),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cgroup_cpus():
    """CPUs the cgroup's quota allows (cpu.max), or None without a quota."""
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            return max(1, int(int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return None


def host_threads() -> int:
    """Every core this process may use: its CPU affinity, bounded by the
    cgroup quota (OMP_NUM_THREADS overrides)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    n = len(os.sched_getaffinity(0))
    q = cgroup_cpus()
    return min(n, q) if q else n


def cpu_info() -> dict:
    """lscpu's model, sockets, cores and threads, plus what this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets = int(info.get("Socket(s)", "1") or 1)
    cps = int(info.get("Core(s) per socket", "0") or 0)
    tpc = int(info.get("Thread(s) per core", "1") or 1)
    return {"model": info.get("Model name", cpu_model()), "sockets": sockets, "cores": sockets * cps,
            "threads": sockets * cps * tpc, "online_cpus": info.get("CPU(s)"),
            "affinity_cpus": len(os.sched_getaffinity(0)), "cgroup_cpu_quota": cgroup_cpus(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "avx512": has_avx512()}


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

KERNEL_OF_LEG = {"exact": "k_search", "fast": "k_search_fast"}


def pmc_profile_path(workload, ef, leg="exact"):
    return os.path.join(PROFILES, f"{KERNEL_OF_LEG[leg]}_{workload}_ef{ef}_pmc.md")


def profiled_traffic(workload, n, nq, k, ef, leg="exact"):
    """HBM bytes per launch of the leg's kernel and its L2 hit rate, recorded
    by a --summarize-pmc run of the same workload, size and ef (None if none)."""
    path = pmc_profile_path(workload, ef, leg)
    if not os.path.exists(path):
        return None, None, None
    for line in open(path):
        if line.startswith("traffic-json: "):
            rec = json.loads(line[len("traffic-json: "):])
            if (rec["n"], rec["nq"], rec["k"], rec["ef"]) == (n, nq, k, ef):
                return rec["traffic_bytes"], os.path.relpath(path, ROOT), rec.get("l2_hit_rate")
    return None, None, None


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
    for key in ("k_search_fast", "k_search", "k_brute_imma", "k_brute_dmma", "k_brute_rowconst", "k_brute_rownorm", "k_brute_merge",
                "k_brute_fill",
                "k_copy16", "k_gather_rows", "k_fill_i32", "k_insert_search", "k_insert_rows", "k_insert_edges",
                "k_rerank", "k_quantize"):
        if key in name:
            return key + (name[name.index("<"):name.index(">") + 1] if "<" in name and key != "k_copy16" else "")
    return name[:60]


def _is_kernel(short, kernel):
    """short names carry template arguments: k_search<...> is not k_search_fast<...>."""
    return short == kernel or short.startswith(kernel + "<")


def _counter_per_dispatch(directory, counter, last=None, kernel="k_search"):
    """Counter values per kernel, in dispatch order; with `last`, only the
    last `last` dispatches of `kernel` (the timed region of a bench run)."""
    rows = []
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r["Counter_Name"] == counter:
            rows.append((int(r["Dispatch_Id"]), _short_kernel(r["Kernel_Name"]), float(r["Counter_Value"])))
    rows.sort()
    out = {}
    for _, key, v in rows:
        out.setdefault(key, []).append(v)
    if last:
        for key in out:
            if _is_kernel(key, kernel):
                out[key] = out[key][-last:]
    return out


def _timed_search_ms(directory, last, kernel="k_search"):
    """Average duration of the last `last` dispatches of `kernel` in a kernel trace."""
    rows = [r for r in _csv_rows(directory, "kernel_trace.csv") if _is_kernel(_short_kernel(r["Kernel_Name"]), kernel)]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    sel = rows[-last:]
    return sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in sel) / len(sel) / 1e6


def summarize(args):
    """Writes a profiles/ summary from rocprofv3 output and the bench line of
    the same command."""
    bl = json.loads(open(args.bench_json).read().strip().splitlines()[-1])
    cfg, roof = bl["config"], bl["roofline"]
    kern = roof.get("kernel", "k_search")
    head = [f"# {kern} — {cfg['workload']}", "",
            f"ef={cfg['ef']}, k={cfg['k']}, {cfg['nq']} queries per launch, recall@10 = {cfg['recall@10']}", "",
            "Commands (committed files only):", ""]
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
            if _is_kernel(name, kern):
                search_avg = avg
            body.append(f"| `{name}` | {r['Calls']} | {avg:.4f} ms | {float(r['MinNs']) / 1e6:.4f} ms | "
                        f"{float(r['MaxNs']) / 1e6:.4f} ms |")
        timed_ms = _timed_search_ms(args.summarize_trace, bl["steps"], kern)
        apq = roof.get("algorithmic_per_query", roof.get("per_query"))
        body += ["", f"The last {bl['steps']} {kern} dispatches are the timed region: {timed_ms:.4f} ms each on "
                 f"average in the trace (the others are the ef sweeps, the counter pass and the autotuner's trials).",
                 f"bench.py in the same run: {roof['kernel_ms']} ms per {kern} launch (HIP events on the "
                 f"index stream), {bl['value']:,.0f} QPS, algorithmic "
                 f"{roof['algorithmic_bytes_per_launch'] / 1e9:.3f} GB per launch -> {roof['achieved']} GB/s "
                 f"= {roof['frac'] * 100:.2f} % of 8 TB/s.",
                 f"Algorithmic work per query (reference Profiler units): {apq['hops']} hops, {apq['neighbors']} "
                 f"adjacency entries scanned, {apq['dists']} distances."]
        if "kernel_per_query" in roof:
            kq = roof["kernel_per_query"]
            why = ("ids its LDS visited filter forgot are read again" if kern == "k_search_fast"
                   else "the reference's own counts: its visited set is exact")
            body += [f"The kernel's own counters per query: {kq['hops']} hops, {kq['dists']} distances "
                     f"({why}): {roof['kernel_counted_bytes_per_launch'] / 1e9:.3f} GB per launch."]
    if args.summarize_pmc:
        fdir, wdir = args.summarize_pmc
        fetch = _counter_per_dispatch(fdir, "FETCH_SIZE", last=bl["steps"], kernel=kern)
        write = _counter_per_dispatch(wdir, "WRITE_SIZE", last=bl["steps"], kernel=kern)
        kib = CALIB_BYTES / 1024.0
        ff = (sum(fetch["k_copy16"]) / len(fetch["k_copy16"])) / kib
        fw = (sum(write["k_copy16"]) / len(write["k_copy16"])) / kib
        skey = [key for key in fetch if _is_kernel(key, kern)][0]
        f_kib = sum(fetch[skey]) / len(fetch[skey])
        w_kib = sum(write[skey]) / len(write[skey])
        fb, wb = f_kib * 1024 / ff, w_kib * 1024 / fw
        alg = roof["algorithmic_bytes_per_launch"]
        body += ["", "Hardware byte counters (two `--pmc` passes, KiB per dispatch, averaged over the "
                 f"dispatches of each kernel; for {kern} the timed region's last dispatches):", "",
                 "| Kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB |", "|---|---|---|---|"]
        for key in sorted(set(fetch) | set(write)):
            fv, wv = fetch.get(key, [0]), write.get(key, [0])
            body.append(f"| `{key}` | {len(fv)} | {sum(fv) / len(fv):,.1f} | {sum(wv) / len(wv):,.1f} |")
        body += ["", f"Calibration: `k_copy16` reads and writes {CALIB_BYTES >> 20} MiB "
                 f"({kib:,.0f} KiB); FETCH_SIZE reads {ff:.4f} x and WRITE_SIZE {fw:.4f} x that.",
                 f"{kern} per launch: {fb / 1e9:.3f} GB fetched + {wb / 1e9:.3f} GB written = "
                 f"{(fb + wb) / 1e9:.3f} GB of HBM traffic, against {alg / 1e9:.3f} GB algorithmic "
                 f"({(fb + wb) / alg:.2f} x).", ""]
        rec = {"kernel": kern, "workload": args.workload, "n": cfg["n"], "nq": cfg["nq"], "k": cfg["k"], "ef": cfg["ef"],
               "fetch_bytes": int(fb), "write_bytes": int(wb), "traffic_bytes": int(fb + wb),
               "algorithmic_bytes": alg, "factor_fetch": round(ff, 5), "factor_write": round(fw, 5)}
        if args.summarize_l2:
            hits = _counter_per_dispatch(args.summarize_l2, "TCC_HIT_sum", last=bl["steps"], kernel=kern)
            miss = _counter_per_dispatch(args.summarize_l2, "TCC_MISS_sum", last=bl["steps"], kernel=kern)
            body += ["L2 (TCC) requests, a third `--pmc` pass (TCC_HIT_sum, TCC_MISS_sum per dispatch):", "",
                     "| Kernel | hits | misses | hit rate |", "|---|---|---|---|"]
            for key in sorted(set(hits) & set(miss)):
                hv, mv = sum(hits[key]) / len(hits[key]), sum(miss[key]) / len(miss[key])
                rate = hv / (hv + mv) if hv + mv else 0.0
                body.append(f"| `{key}` | {hv:,.0f} | {mv:,.0f} | {rate * 100:.1f} % |")
            hk = [key for key in hits if _is_kernel(key, kern)][0]
            hv, mv = sum(hits[hk]) / len(hits[hk]), sum(miss[hk]) / len(miss[hk])
            rec["l2_hit_rate"] = round(hv / (hv + mv), 4)
            body += ["", f"{kern} L2 hit rate: {rec['l2_hit_rate'] * 100:.1f} % of its L2 requests.", ""]
        body.append("traffic-json: " + json.dumps(rec))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(head + body) + "\n")
    print(f"wrote {args.out}")


def run_serving(args):
    """Open-loop serving runs: Poisson arrivals of query batches (sizes
    log-uniform in [1, 4096], ef uniform in [16, 256], k in {10, 100}) mixed
    with insert batches of 1,024 new vectors at each --insert-share (a comma
    list; the mixes run one after another on one GPU-built index, inserts
    accumulating), through the library's serving loop.  Reports p50 / p99
    request latency, completed queries per second and the loop's batching
    for every mix, in one JSON line (the first mix is its value)."""
    import threading

    from pickle_amd import abi, datasets

    wl = WORKLOADS[args.workload]
    n = args.n or wl["n"]
    n_req = args.requests
    shares = [float(x) for x in str(args.insert_share).split(",")]
    n_ins_all = [int(round(n_req * sh)) if sh > 0 else 0 for sh in shares]
    n_ins_all = [max(1, x) if sh > 0 else 0 for x, sh in zip(n_ins_all, shares)]
    n_ranges = 2 if args.target_recall else 1  # the mixes run once per ef range
    extra = n_ranges * sum(n_ins_all) * 1024
    streamed = n + extra >= STREAM_ROWS
    if streamed:
        # rows are generated on a second thread while earlier ones are uploaded
        # and inserted (the generator releases the interpreter lock)
        st, params = datasets.workload_stream(args.workload, n=n + extra, nq=args.nq or 20000)
        dim, dtype = st.centers.shape[1], np.dtype(st.args["out_dtype"])
        base = None
    else:
        base, queries, params = datasets.workload(args.workload, n=n + extra, nq=args.nq or 20000)
        dim, dtype = base.shape[1], base.dtype
    mcode = abi.L2 if params["metric"] == "l2" else abi.IP
    idx = abi.Index(dtype, mcode, dim, params["M"], params["efc"], n + extra)
    if args.row_limit:
        idx.set_row_limit(args.row_limit)
    levels = abi.sample_levels(n + extra, params["M"], args.seed)
    tb = time.perf_counter()
    if streamed:
        import queue

        chunks = queue.Queue(maxsize=2)

        def generate():
            for c in st.base_chunks():
                chunks.put(c)
            chunks.put(None)

        tg = threading.Thread(target=generate, daemon=True)
        tg.start()
        tail, at, started = [], 0, False
        while True:
            c = chunks.get()
            if c is None:
                break
            m = c.shape[0]
            nb = max(0, min(m, n - at))
            if nb:
                idx.set_base(c[:nb], first=at)
                if not started:
                    idx.init_build(levels)
                    started = True
                idx.insert(np.arange(at, at + nb, dtype=np.int32), max_batch=args.insert_batch,
                           exact_rows=not args.capped_rows, graph=not args.stream_build)
            if nb < m:
                tail.append(c[nb:])
            at += m
            log(f"  inserted {min(at, n):,} / {n:,} ({time.perf_counter() - tb:.1f}s)")
        tg.join()
        queries = st.queries()
        extra_rows = np.concatenate(tail) if tail else np.zeros((0, dim), dtype)
    else:
        idx.set_base(base[:n])
        idx.init_build(levels)
        chunk = 1 << 20
        for s0 in range(0, n, chunk):
            idx.insert(np.arange(s0, min(n, s0 + chunk), dtype=np.int32), max_batch=args.insert_batch,
                       exact_rows=not args.capped_rows, graph=not args.stream_build)
            if n > chunk:
                log(f"  inserted {min(n, s0 + chunk):,} / {n:,} ({time.perf_counter() - tb:.1f}s)")
        extra_rows = base[n:]
    build_s = time.perf_counter() - tb
    rows_desc = ("rows capped at D" if args.capped_rows else
                 f"reference rows up to {args.row_limit or 511} ids")
    log(f"serving index: {n} x {dim} {dtype} built on the GPU in {build_s:.1f}s (+{extra} rows reserved, "
        f"{rows_desc})")

    # recall of the served index: exact ground truth on the matrix cores
    # (int8 MFMA for u8 rows) for 1,000 queries, recall@10 by ef
    tg0 = time.perf_counter()
    gq = queries[:1000]
    gt_ids = idx.brute_force(gq)[0][:, :10]
    gt_s = time.perf_counter() - tg0
    search = idx.search_fast if args.serve_fast else idx.search  # the mode the loop serves with
    ef_cap = 512 if args.serve_fast else 630
    recall_by_ef = {}
    for ef in (32, 64, 128, 256, 384, 512, 630):
        if ef > ef_cap:
            continue
        ids = search(gq, 10, ef)[0]
        recall_by_ef[ef] = round(recall_at_k(ids, gt_ids, 10), 4)
    log(f"recall@10 by ef {recall_by_ef} (ground truth for {len(gq)} queries in {gt_s:.2f}s)")

    # ef ranges of the mixes: the workload's [16, 256], and with
    # --target-recall one fixed ef, the smallest whose recall@10 reaches it
    ranges = [(16, 256)]
    ef_target = None
    if args.target_recall:
        lo, hi = 1, ef_cap
        for ef in sorted(recall_by_ef):
            if recall_by_ef[ef] >= args.target_recall:
                hi = ef
                break
            lo = ef
        while hi - lo > max(4, hi // 32):  # bisect between the measured points
            mid = (lo + hi) // 2
            r = recall_at_k(search(gq, 10, mid)[0], gt_ids, 10)
            recall_by_ef[mid] = round(r, 4)
            lo, hi = (mid, hi) if r < args.target_recall else (lo, mid)
        ef_target = hi
        ranges.append((ef_target, ef_target))
        log(f"ef {ef_target}: recall@10 {recall_by_ef.get(ef_target)} >= {args.target_recall}")
    max_ef = max(r[1] for r in ranges)

    # capacity probe per range: one slot-sized launch at the range's mean ef
    d_q = abi.DeviceBuffer.from_array(queries[:4096])
    d_i, d_d, d_c = abi.DeviceBuffer(4096 * 100 * 4), abi.DeviceBuffer(4096 * 100 * 4), abi.DeviceBuffer(4096 * 4)
    mean_size = 4095 / np.log(4096)
    rates = {}
    for lo, hi in ranges:
        ef_mean = (lo + hi) // 2
        for _ in range(2):
            idx.timer_begin()
            (idx.search_fast_device if args.serve_fast else idx.search_device)(d_q, 4096, 10, ef_mean, 10, d_i, d_d,
                                                                                d_c, None)
            cap_ms = idx.timer_end()
        cap_qps = 4096 / (cap_ms / 1e3)
        rates[(lo, hi)] = (args.load * cap_qps / mean_size, cap_qps, ef_mean)
        log(f"capacity ~{cap_qps:,.0f} QPS at ef={ef_mean}; offered load {args.load:.2f} -> "
            f"{rates[(lo, hi)][0]:,.1f} requests/s")

    rng = np.random.default_rng(args.seed)
    ins_at = n
    mixes = []
    slog = []
    plan = [(r, sh, ni) for r in ranges for sh, ni in zip(shares, n_ins_all)]
    for mix, ((ef_lo, ef_hi), share, n_ins) in enumerate(plan):
        rate, cap_qps, ef_mean = rates[(ef_lo, ef_hi)]
        kinds = rng.permutation(np.r_[np.ones(n_ins, bool), np.zeros(n_req - n_ins, bool)])
        gaps = rng.exponential(1.0 / rate, n_req)
        sizes = np.minimum(4096, np.floor(np.exp(rng.uniform(0, np.log(4096), n_req)))).astype(int)
        efs = rng.integers(ef_lo, ef_hi + 1, n_req)
        ks = rng.choice([10, 100], n_req)
        efs = np.maximum(efs, ks)
        srv = abi.Server(idx, slots=args.slots, slot_queries=4096, max_k=100, max_ef=max(max_ef, 100),
                         fast=args.serve_fast)
        tickets = []
        done = []
        t0 = time.perf_counter()

        def producer(ins_first):
            t_next = t0
            q_at, at = 0, ins_first
            for i in range(n_req):
                t_next += gaps[i]
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                if kinds[i]:
                    ids = np.arange(at, at + 1024, dtype=np.int32)
                    tickets.append(("insert", 1024, srv.insert(ids, extra_rows[at - n:at - n + 1024])))
                    at += 1024
                else:
                    m = int(sizes[i])
                    s0 = q_at % (len(queries) - m)
                    tickets.append(("search", m, srv.submit(queries[s0:s0 + m], int(ks[i]), int(efs[i]))))
                    q_at += m
            tickets.append(None)

        th = threading.Thread(target=producer, args=(ins_at,))
        th.start()
        j = 0
        while True:
            while j >= len(tickets):
                time.sleep(0.0005)
            t = tickets[j]
            j += 1
            if t is None:
                break
            kind, m, tk = t
            _, _, _, lat, que = srv.wait(tk)
            done.append((kind, m, lat, que))
            if args.serve_log:
                slog.append({"mix": mix, "i": j - 1, "kind": kind, "size": m, "latency_us": round(lat, 1),
                             "queue_us": round(que, 1), "done_s": round(time.perf_counter() - t0, 6)})
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        ins_at += n_ins * 1024
        lat_s = np.array([d[2] for d in done if d[0] == "search"]) / 1e3
        lat_i = np.array([d[2] for d in done if d[0] == "insert"]) / 1e3
        nq_done = sum(d[1] for d in done if d[0] == "search")
        res = {"ef": f"uniform [{ef_lo}, {ef_hi}]" if ef_lo < ef_hi else ef_lo,
               "recall@10_at_ef": recall_by_ef.get(ef_lo) if ef_lo == ef_hi else None,
               "offered_requests_per_s": round(rate, 1), "capacity_qps_probe": round(cap_qps, 1),
               "insert_share": share, "requests": n_req, "inserts": n_ins, "qps": round(nq_done / wall, 1),
               "p50_request_latency_ms": round(float(np.percentile(lat_s, 50)), 3),
               "p99_request_latency_ms": round(float(np.percentile(lat_s, 99)), 3),
               "p50_insert_latency_ms": round(float(np.percentile(lat_i, 50)), 3) if len(lat_i) else None,
               "queries": int(nq_done), "wall_s": round(wall, 3), "rows_after": int(ins_at), "loop": st}
        mixes.append(res)
        log(f"mix {mix} (ef {res['ef']}, insert share {share}): served {nq_done} queries in {n_req - n_ins} requests + {n_ins} "
            f"inserts in {wall:.2f}s: {res['qps']:,.0f} QPS, request latency p50 {res['p50_request_latency_ms']} ms, "
            f"p99 {res['p99_request_latency_ms']} ms; loop {st}")
    if args.serve_log:
        with open(args.serve_log, "w") as f:
            for rec in slog:
                f.write(json.dumps(rec) + "\n")
    first = mixes[0]
    line = {
        "metric": "serving: p50 / p99 request latency and completed QPS under open-loop Poisson arrivals",
        "value": first["qps"], "unit": "queries/s", "n_gpus": 1, "steps": n_req, "warmup": 0,
        "ms_per_step": round(first["wall_s"] / n_req * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": str(dtype), "data": "synthetic",
        "config": {"workload": f"{wl['desc']} (n={n}, GPU-built, +{extra} streamed)", "requests": n_req,
                   "insert_batch": 1024, "slots": args.slots, "offered_load": args.load,
                   "arrival_rate_per_s": first["offered_requests_per_s"], "request_sizes": "log-uniform [1, 4096]",
                   "ef": "uniform [16, 256]" + (f"; then fixed at {ef_target}, the smallest ef with recall@10 >= "
                                                f"{args.target_recall}" if ef_target else ""), "k": "{10, 100}",
                   "p50_request_latency_ms": first["p50_request_latency_ms"],
                   "p99_request_latency_ms": first["p99_request_latency_ms"],
                   "mixes": mixes, "build_seconds": round(build_s, 2), "rows": rows_desc,
                   "search": "fast mode (PK_SERVE_FAST)" if args.serve_fast else "exact (reference heap order)",
                   "recall@10_by_ef": recall_by_ef, "ground_truth_seconds": round(gt_s, 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def host_ground_truth(base, queries, k, metric):
    """Exact top-k on the host: FP32 GEMM blocks, top 2k candidates, FP64
    recheck (for k larger than the GPU brute force keeps)."""
    B = np.asarray(base, np.float32)
    bn = (B.astype(np.float64) ** 2).sum(1) if metric == "l2" else None
    out = np.empty((len(queries), k), np.int64)
    for s0 in range(0, len(queries), 256):
        Q = np.asarray(queries[s0:s0 + 256], np.float32)
        dots = Q @ B.T
        d = (bn[None, :] - 2 * dots) if metric == "l2" else -dots
        cand = np.argpartition(d, 2 * k, axis=1)[:, :2 * k]
        for i in range(len(Q)):
            c = cand[i]
            A, q = B[c].astype(np.float64), Q[i].astype(np.float64)
            d64 = ((A - q) ** 2).sum(1) if metric == "l2" else -(A @ q)
            out[s0 + i] = c[np.lexsort((c, d64))][:k]
    return out


def run_uneven(args):
    """Uneven-batch persistent search on a GPU-built graph: 1M x 96 unit
    vectors (vMF-like mixture, IP), the GPU batched build compared with the
    reference's CPU build by recall vs ef, then one launch of all queries with
    per-query ef log-uniform in [16, 512] and k in {1, 10, 100}, autotuned
    and captured in a HIP graph; the timed steps replay the graph."""
    from pickle_amd import abi, datasets

    wl = WORKLOADS[args.workload]
    n, nq = args.n or wl["n"], args.nq or wl["nq"]
    base, queries, params = datasets.workload(args.workload, n=n, nq=nq)
    dim, metric, M, efc = base.shape[1], params["metric"], params["M"], params["efc"]
    mcode = abi.L2 if metric == "l2" else abi.IP
    threads = host_threads()
    binary, isa = ref_binary()
    work = args.graph_cache or tempfile.mkdtemp(prefix="pk_bench_")
    os.makedirs(work, exist_ok=True)
    base_path, graph_path = os.path.join(work, "uneven.base"), os.path.join(work, "uneven_cpu.graph")
    base.tofile(base_path)
    # graphs: GPU batched build and the reference's CPU build
    idx = abi.Index(base.dtype, mcode, dim, M, efc, n)
    tb = time.perf_counter()
    idx.build(base, max_batch=args.insert_batch, seed=args.seed, graph=not args.stream_build,
              progress=lambda d, t: log(f"  GPU build {d:,} / {t:,} ({time.perf_counter() - tb:.1f}s)"))
    idx.sync()
    gpu_build_s = time.perf_counter() - tb
    log(f"GPU batched build in {gpu_build_s:.1f}s")
    cpu_info = cpu_build(binary, base_path, n, dim, "f32", metric, M, efc, threads, graph_path)
    log(f"reference CPU build ({isa}, {threads} threads) in {cpu_info['build_seconds']:.1f}s")
    cidx = abi.Index(base.dtype, mcode, dim, M, efc, n)
    cidx.set_base(base)
    cidx.load_file(graph_path)

    # exact ground truth (top 100) on a query subset
    ns = min(nq, args.gt_queries or 1000)
    t1 = time.time()
    gt = host_ground_truth(base, queries[:ns], 100, metric)
    log(f"host ground truth (top 100, {ns} queries) in {time.time() - t1:.1f}s")
    curve = []
    for ef in (16, 32, 64, 128, 256, 512):
        rg = recall_at_k(idx.search(queries[:ns], 10, ef)[0], gt, 10)
        rc = recall_at_k(cidx.search(queries[:ns], 10, ef)[0], gt, 10)
        curve.append([ef, round(rg, 4), round(rc, 4)])
        log(f"recall@10 ef={ef}: GPU-built {rg:.4f}, reference-built {rc:.4f}")
    cidx.close()

    # the uneven launch
    rng = np.random.default_rng(args.seed)
    efs = np.exp(rng.uniform(np.log(16), np.log(512), nq)).astype(np.int32)
    ks = rng.choice([1, 10, 100], nq).astype(np.int32)
    efs = np.maximum(efs, ks)
    d_q = abi.DeviceBuffer.from_array(queries)
    d_k, d_e = abi.DeviceBuffer.from_array(ks), abi.DeviceBuffer.from_array(efs)
    d_ids, d_d = abi.DeviceBuffer(nq * 100 * 4), abi.DeviceBuffer(nq * 100 * 4)
    d_c = abi.DeviceBuffer(nq * 4)
    tuned = None
    if not args.no_autotune:
        tuned = idx.autotune(d_q, min(nq, 4096), 10, int(np.median(efs)))
        log(f"autotuned at the median ef {int(np.median(efs))}: {tuned}")
    graph = idx.search_graph(d_q, nq, d_k, d_e, 100, d_ids, d_d, d_c, ef_max=int(efs.max()))
    for _ in range(args.warmup):
        graph.launch()
    idx.sync()
    tw0 = time.perf_counter()
    idx.timer_begin()
    for _ in range(args.steps):
        graph.launch()
    ev_ms = idx.timer_end()
    wall = time.perf_counter() - tw0
    graph.close()
    qps = nq * args.steps / wall
    ids = d_ids.to_array(np.int32, (nq, 100))

    def recall_by_k(ids):
        rec = {}
        for kk in (1, 10, 100):
            sel = np.where(ks[:ns] == kk)[0]
            hit = sum(len(set(ids[i, :kk].tolist()) & set(gt[i, :kk].tolist())) for i in sel)
            rec[f"k={kk}"] = round(hit / (kk * max(1, len(sel))), 4)
        return rec

    rec = recall_by_k(ids)
    log(f"uneven launch: {nq} queries, {args.steps} graph replays: {qps:,.0f} QPS, kernel "
        f"{ev_ms / args.steps:.3f} ms; recall by k {rec}")

    # the same launch in the fast mode (per-query k and ef, ef_max 512)
    fast = None
    if int(efs.max()) <= 512:
        if not args.no_autotune:
            idx.autotune_fast(d_q, min(nq, 4096), 10, int(np.median(efs)))
        for _ in range(args.warmup):
            idx.search_fast_device(d_q, nq, 0, 0, 100, d_ids, d_d, d_c, None, d_k, d_e, int(efs.max()))
        idx.sync()
        tf0 = time.perf_counter()
        for _ in range(args.steps):
            idx.search_fast_device(d_q, nq, 0, 0, 100, d_ids, d_d, d_c, None, d_k, d_e, int(efs.max()))
        idx.sync()
        fwall = time.perf_counter() - tf0
        if idx.device_error():
            sys.exit("fast uneven launch: device reported a failure")
        frec = recall_by_k(d_ids.to_array(np.int32, (nq, 100)))
        fast = {"qps": round(nq * args.steps / fwall, 1), "recall_by_k": frec}
        log(f"fast mode, same launch: {fast['qps']:,.0f} QPS; recall by k {frec}")
    line = {
        "metric": "QPS of one uneven launch (per-query ef log-uniform [16, 512], k in {1, 10, 100}), HIP graph",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl["desc"], "n": n, "nq": nq, "dim": dim, "M": M, "efC": efc,
                   "ef": "log-uniform [16, 512] per query", "k": "{1, 10, 100} per query",
                   "recall_by_k": rec, "fast_mode": fast, "recall_queries": ns,
                   "gpu_build_seconds": round(gpu_build_s, 2),
                   "reference_cpu_build_seconds": round(cpu_info["build_seconds"], 2), "reference_build_threads":
                   threads, "recall10_vs_ef_gpu_built_vs_reference_built": curve, "autotuned": tuned,
                   "kernel_ms": round(ev_ms / args.steps, 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def library_vbits(ef, n):
    """The library's visited-table size for an ef (index.hip,
    default_vbits): by ef, at most twice the capacity."""
    v = 14 if ef <= 32 else 15 if ef <= 64 else 16 if ef <= 256 else 17 if ef <= 512 else 18
    cap_bits = 10
    while cap_bits < 20 and (1 << cap_bits) < 2 * n:
        cap_bits += 1
    return min(v, cap_bits)


GRID_K = (1, 10, 100)  # benchmark/bench_util.hpp:24
GRID_EF = (1, 5, 10, 20, 30, 50, 70, 90, 100, 200, 300, 400, 500)  # benchmark/bench_util.hpp:25


def run_grid(args):
    """The reference's own sweep (benchmark/hnsw_pickle.cpp:64-121): every k
    in {1, 10, 100} against every ef of bench_util.hpp:25, ef < k skipped, on
    one reference-built graph.  Per (k, ef): recall (%, label matches), the
    GPU exact and fast searches' QPS over the whole query batch (inputs and
    outputs in HBM, HIP events), and the reference AnnSearch's QPS and QPS per
    thread (qps/t, the reference's own figure) on the host cores, with its
    Profiler hops and distances per query."""
    from pickle_amd import abi, datasets

    wl = WORKLOADS[args.workload]
    n, nq = args.n or wl["n"], args.nq or wl["nq"]
    base, queries, params = datasets.workload(args.workload, n=n, nq=nq)
    dim, metric, M, efc = base.shape[1], params["metric"], params["M"], params["efc"]
    dtype_name = {np.dtype(np.uint8): "u8", np.dtype(np.int8): "i8", np.dtype(np.float32): "f32",
                  np.dtype(np.float16): "f16"}[base.dtype]
    threads = host_threads()
    binary, isa = ref_binary()
    if binary is None:
        sys.exit("the reference build (oracle/_ref) is missing")
    work = args.graph_cache or tempfile.mkdtemp(prefix="pk_grid_")
    os.makedirs(work, exist_ok=True)
    base_path = os.path.join(work, f"grid_{args.workload}_n{n}.base")
    graph_path = os.path.join(work, f"grid_{args.workload}_n{n}.graph")
    base.tofile(base_path)
    log(f"grid on {args.workload}: n={n} nq={nq}; reference CPU build ({isa}, {threads} threads) ...")
    binfo = cpu_build(binary, base_path, n, dim, dtype_name, metric, M, efc, threads, graph_path)
    log(f"built in {binfo['build_seconds']:.1f}s")
    idx = abi.Index(base.dtype, abi.L2 if metric == "l2" else abi.IP, dim, M, efc, n)
    idx.set_base(base)
    idx.load_file(graph_path)

    ns = min(nq, args.gt_queries or 1000)
    t1 = time.time()
    gt = host_ground_truth(base, queries[:ns], max(GRID_K), metric)
    log(f"host ground truth (top {max(GRID_K)}, {ns} queries) in {time.time() - t1:.1f}s")
    cnq = min(nq, args.cpu_queries or 2000)
    q_path = os.path.join(work, f"grid_q{cnq}.queries")
    queries[:cnq].tofile(q_path)

    d_q = abi.DeviceBuffer.from_array(queries)
    kmax = max(GRID_K)
    d_ids, d_d, d_c = abi.DeviceBuffer(nq * kmax * 4), abi.DeviceBuffer(nq * kmax * 4), abi.DeviceBuffer(nq * 4)

    def gpu_leg(fast, k, ef):
        launch = idx.search_fast_device if fast else idx.search_device
        while True:
            launch(d_q, nq, k, ef, k, d_ids, d_d, d_c)
            idx.sync()
            idx.timer_begin()
            for _ in range(3):
                launch(d_q, nq, k, ef, k, d_ids, d_d, d_c)
            ms = idx.timer_end() / 3
            flags = idx.device_error()
            if not fast and flags == 1 and max(vb[0], library_vbits(ef, n)) < 20:
                # a visited table filled: the next size per wave
                vb[0] = max(vb[0], library_vbits(ef, n)) + 1
                idx.set_tuning(0, vb[0], 0)
                continue
            break
        if flags:
            sys.exit(f"grid: device search failure flags {flags} at k={k} ef={ef}")
        ids = d_ids.to_array(np.int32, (nq, k))[:ns]
        return round(100.0 * recall_at_k(ids, gt, k), 2), round(nq / (ms / 1e3), 1)

    rows = []
    for k in GRID_K:
        efs = [ef for ef in GRID_EF if ef >= k]  # hnsw_pickle.cpp:74
        vb = [0]
        idx.set_tuning(0, 0, 0)
        out = os.path.join(work, f"grid_cpu_k{k}")
        st = cpu_search_counters(binary, graph_path, base_path, q_path, n, cnq, dim, dtype_name, k, efs, threads,
                                 out)
        for ef, cs in zip(efs, st):
            cids = read_ref_results(f"{out}.ef{ef}", cnq, k)
            r_n = min(cnq, ns)
            c_rec = round(100.0 * recall_at_k(cids[:r_n], gt[:r_n], k), 2)
            x_rec, x_qps = gpu_leg(False, k, ef)
            same = float(np.mean(np.all(d_ids.to_array(np.int32, (nq, k))[:cnq] == cids, axis=1)))
            f_rec, f_qps = gpu_leg(True, k, ef)
            row = {"k": k, "ef": ef, "gpu_exact": {"recall": x_rec, "qps": x_qps, "ids_equal_reference": round(same, 4)},
                   "gpu_fast": {"recall": f_rec, "qps": f_qps},
                   "cpu_reference": {"recall": c_rec, "qps": round(cs["qps"], 1),
                                     "qps_per_thread": round(cs["qps"] / threads, 1),
                                     "hop": round(cs["hops"] / cnq, 1), "dist": round(cs["dists"] / cnq, 1),
                                     "base_hop": round(cs["hops0"] / cnq, 1),
                                     "base_dist": round(cs["dists0"] / cnq, 1)}}
            rows.append(row)
            log(f"k={k:3d} ef={ef:3d} | CPU recall={c_rec:6.2f} qps={cs['qps']:10,.0f} qps/t={cs['qps'] / threads:8,.0f} "
                f"hop={cs['hops'] / cnq:6.1f} dist={cs['dists'] / cnq:8.1f} | GPU exact recall={x_rec:6.2f} "
                f"qps={x_qps:12,.0f} ids=ref {same * 100:.1f}% | GPU fast recall={f_rec:6.2f} qps={f_qps:12,.0f}")
    hostcpu = cpu_info()
    line = {"metric": "recall and QPS over the reference's k x ef grid (benchmark/hnsw_pickle.cpp:73-121)",
            "value": None, "unit": "queries/s", "n_gpus": 1, "steps": 3, "warmup": 1, "higher_is_better": True,
            "dtype": dtype_name, "data": "synthetic",
            "config": {"workload": wl["desc"], "n": n, "nq": nq, "recall_queries": ns, "cpu_queries": cnq,
                       "graph": f"reference CPU build ({binfo['build_seconds']:.1f}s, {threads} threads)",
                       "cpu": {"threads": threads, "isa": isa, "host": hostcpu}},
            "grid": rows}
    print(json.dumps(line), flush=True)


def cpu_search_counters(binary, graph, base_path, q_path, n, nq, dim, dtype, k, efs, threads, out):
    """The reference harness's search over a list of ef values: wall time of
    one timed pass each, then an untimed pass with the Profiler on
    (benchmark/hnsw_pickle.cpp:79-92)."""
    cmd = [binary, "search", "--graph", graph, "--base", base_path, "--queries", q_path, "--n", str(n),
           "--nq", str(nq), "--dim", str(dim), "--dtype", dtype, "--k", str(k), "--ef", ",".join(map(str, efs)),
           "--threads", str(threads), "--repeat", "1", "--counters", "1", "--out", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference search failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout)


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
    ap.add_argument("--workload", default="gm10m_f32", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="override the base size (smaller runs)")
    ap.add_argument("--nq", type=int, default=None)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--recall-target", type=float, default=0.95)
    ap.add_argument("--efs", default=None, help="ef sweep (default per workload)")
    ap.add_argument("--mode", choices=["fast", "exact"], default="fast",
                    help="the headline leg: fast mode or the exact (reference heap order) search")
    ap.add_argument("--ef", type=int, default=0, help="fixed ef (skips the sweep)")
    ap.add_argument("--graph-cache", default=None, help="directory to reuse / keep the CPU-built graph")
    ap.add_argument("--no-cpu-search", action="store_true", help="skip the CPU search baseline")
    ap.add_argument("--cpu-repeat", type=int, default=3)
    ap.add_argument("--cpu-queries", type=int, default=None,
                    help="bound the CPU search sample (0 = all queries; default per workload)")
    ap.add_argument("--gt-queries", type=int, default=None, help="queries with ground truth (0 = all; default per "
                                                                 "workload)")
    ap.add_argument("--build", choices=["cpu", "gpu"], default=None,
                    help="graph from the reference's CPU build or the GPU batched build (default per workload)")
    ap.add_argument("--insert-batch", type=int, default=16384)
    ap.add_argument("--stream-build", action="store_true",
                    help="GPU builds launch each insert batch's kernels directly instead of replaying them as HIP "
                         "graphs (PK_BUILD_GRAPH, the default)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", default=None,
                    help="comma list of low-precision traversal copies (f16, i8) searched + FP32 re-ranked "
                         "beside the FP32 legs (float32 workloads), or none")
    ap.add_argument("--no-autotune", action="store_true", help="time with the default tuning")
    ap.add_argument("--fast-tuning", default=None, metavar="W,HB,WPG,BPC",
                    help="fixed fast-mode tuning (width, hash bits, waves per group, workgroups per CU; 0 = "
                         "automatic) instead of the autotuner's, e.g. the tuning a profiled run chose")
    ap.add_argument("--base-file", default=None, help=".npy / .fbin / .u8bin / .i8bin base vectors instead of "
                                                      "a synthetic workload (with --query-file)")
    ap.add_argument("--query-file", default=None)
    ap.add_argument("--metric", choices=["l2", "ip"], default="l2", help="with --base-file")
    ap.add_argument("--M", type=int, default=32, help="with --base-file")
    ap.add_argument("--efc", type=int, default=None, help="efConstruction (with --base-file: 200 by default)")
    ap.add_argument("--serve", action="store_true", help="open-loop serving run through the request loop")
    ap.add_argument("--grid", action="store_true",
                    help="the reference's k x ef grid (k 1/10/100, ef 1..500) on a reference-built graph")
    ap.add_argument("--uneven", action="store_true",
                    help="uneven-batch launch (per-query ef / k) on a GPU-built graph, HIP graph replay")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--row-limit", type=int, default=None,
                    help="GPU builds: longest row the reference-rows insertion writes (default per workload; 0: 511)")
    ap.add_argument("--capped-rows", action="store_true",
                    help="GPU builds: rows never exceed D ids (the reference's AddEdge can grow a row past D)")
    ap.add_argument("--insert-share", default="0,0.05",
                    help="comma list of insert shares, one serving run each on the same index")
    ap.add_argument("--load", type=float, default=0.3, help="offered search load as a fraction of measured capacity")
    ap.add_argument("--serve-fast", action="store_true",
                    help="serving: launches run the fast mode (recall measured with it; ef <= 512)")
    ap.add_argument("--target-recall", type=float, default=None,
                    help="serving: also run the mixes at the smallest ef whose recall@10 reaches this")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--serve-log", default=None, help="per-request latencies (JSON lines) of a serving run")
    ap.add_argument("--summarize-trace", default=None, metavar="DIR",
                    help="write a profiles/ summary from a rocprofv3 --kernel-trace --stats output directory")
    ap.add_argument("--summarize-l2", default=None, metavar="TCC_DIR",
                    help="with --summarize-pmc: a --pmc TCC_HIT_sum TCC_MISS_sum run's directory")
    ap.add_argument("--summarize-pmc", nargs=2, default=None, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write a profiles/ summary from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE directories")
    ap.add_argument("--bench-json", default=None, help="the JSON line bench.py printed in the profiled run")
    ap.add_argument("--out", default=None, help="summary file to write")
    ap.add_argument("--cmd", action="append", default=[], help="command lines that produced the profile")
    args = ap.parse_args()

    if args.summarize_trace or args.summarize_pmc:
        summarize(args)
        return

    if args.uneven:
        run_uneven(args)
        return

    if args.grid:
        if args.gpus != 1:
            sys.exit("bench.py measures one GPU only")
        run_grid(args)
        return

    if args.serve:
        if args.gpus != 1:
            sys.exit("bench.py measures one GPU only")
        run_serving(args)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (run with --gpus 1). "
              "Sharding across GPUs is a design, not part of this measurement.", file=sys.stderr)
        sys.exit(2)

    from pickle_amd import abi, datasets

    wl = WORKLOADS[args.workload]
    n = args.n or wl["n"]
    nq = args.nq or wl["nq"]
    k = args.k
    if args.efs is None:
        args.efs = wl.get("efs", "10,12,14,16,20,24,28,32,40,48,64,96,128,192,256,384,512")
    if args.gt_queries is None:
        args.gt_queries = wl.get("gt_queries", 0)
    if args.cpu_queries is None:
        args.cpu_queries = wl.get("cpu_queries", 0)
    t0 = time.time()
    if args.base_file:
        from pickle_amd import formats

        base = np.ascontiguousarray(formats.read_vectors(args.base_file, 0, args.n))
        queries = np.ascontiguousarray(formats.read_vectors(args.query_file, 0, args.nq))
        n, nq = base.shape[0], queries.shape[0]
        params = {"metric": args.metric, "M": args.M, "efc": args.efc or 200}
        wl = dict(wl, desc=f"{os.path.basename(args.base_file)} ({n} x {base.shape[1]} {base.dtype}), "
                           f"{nq} queries from {os.path.basename(args.query_file)}, {args.metric}, M={args.M}, "
                           f"efC={args.efc}", n=n, nq=nq)
    else:
        base, queries, params = datasets.workload(args.workload, n=n, nq=nq)
    dtype_name = {np.dtype(np.uint8): "u8", np.dtype(np.int8): "i8", np.dtype(np.float32): "f32",
                  np.dtype(np.float16): "f16"}[base.dtype]
    dim = base.shape[1]
    metric = params["metric"]
    M, efc = params["M"], params["efc"]
    if args.efc and not args.base_file:
        efc = args.efc
    log(f"workload {args.workload}: n={n} nq={nq} dim={dim} {dtype_name} {metric} M={M} efC={efc} "
        f"(generated in {time.time() - t0:.1f}s)")

    threads = host_threads()
    binary, isa = ref_binary()
    if binary is None:
        sys.exit("the reference build (oracle/_ref) is missing: run __graft_entry__.build() where the "
                 "reference sources are")
    build = args.build or wl.get("build", "cpu")
    prec = args.precision or wl.get("precision", "none")
    if prec == "none":
        prec = ""
    if prec and base.dtype != np.float32:
        sys.exit("--precision needs a float32 workload")

    work = args.graph_cache or tempfile.mkdtemp(prefix="pk_bench_")
    os.makedirs(work, exist_ok=True)
    tag = f"{os.path.basename(args.base_file) if args.base_file else args.workload}_n{n}_{build}"
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
        row_limit = args.row_limit if args.row_limit is not None else wl.get("row_limit", 0)
        idx.set_row_limit(row_limit)
        tb = time.perf_counter()
        idx.build(base, max_batch=args.insert_batch, seed=args.seed, exact_rows=not args.capped_rows,
                  graph=not args.stream_build,
                  progress=lambda d, t: log(f"  inserted {d:,} / {t:,} ({time.perf_counter() - tb:.1f}s)"))
        idx.sync()
        build_info = {"build_seconds": time.perf_counter() - tb, "threads": 0, "where": "gpu",
                      "hip_graphs": not args.stream_build,
                      "insert_batch": args.insert_batch, "capped_rows": bool(args.capped_rows), "efc": efc,
                      "row_limit": row_limit or 511}
        log(f"GPU build in {build_info['build_seconds']:.1f}s")
        idx.save_file(graph_path)  # for the reference's search on the same graph
        json.dump(build_info, open(meta_path, "w"))
    info = idx.info()
    log(f"graph in HBM: layers {info['sizes']}")

    gt_nq = min(nq, args.gt_queries) if args.gt_queries else nq
    t1 = time.time()
    parts = []
    for s0 in range(0, gt_nq, 10000):
        parts.append(idx.brute_force(queries[s0:min(gt_nq, s0 + 10000)]))
        if n >= 5_000_000:
            log(f"  ground truth {min(gt_nq, s0 + 10000):,} / {gt_nq:,} ({time.time() - t1:.1f}s)")
    gt_ids = np.concatenate([p_[0] for p_ in parts])
    if base.dtype.kind == "f":
        # recheck: the FP64 GEMM form |x|^2 - 2 q.x + |q|^2 can differ from the
        # direct FP64 sum in its last bits; the 32 candidates are re-ranked by
        # the direct sum
        for i in range(gt_nq):
            c = gt_ids[i]
            A, Q = base[c].astype(np.float64), queries[i].astype(np.float64)
            d64 = ((A - Q) ** 2).sum(1) if metric == "l2" else -(A @ Q)
            gt_ids[i] = c[np.lexsort((c, d64))]
    gt_s = time.time() - t1
    log(f"ground truth (GPU brute force, {gt_nq} queries) in {gt_s:.1f}s")
    # the brute force is a dense contraction on the matrix cores: int8 MFMA
    # for u8/i8 rows, FP64 MFMA for f16/f32 (wall time, host staging included)
    gt_ops = 2.0 * gt_nq * n * dim
    ground_truth = {"queries": gt_nq, "seconds": round(gt_s, 4),
                    "kernel": "k_brute_imma" if base.dtype.kind in "ui" else "k_brute_dmma",
                    "tera_ops_per_s": round(gt_ops / gt_s / 1e12, 2),
                    "mfma_peak_tera_ops_per_s": 5000.0 if base.dtype.kind in "ui" else 78.6}

    # HBM-resident buffers for the timed runs
    d_q = abi.DeviceBuffer.from_array(queries)
    d_ids = abi.DeviceBuffer(nq * k * 4)
    d_dist = abi.DeviceBuffer(nq * k * 4)
    d_cnt = abi.DeviceBuffer(nq * 4)
    d_ctr = abi.DeviceBuffer(nq * abi.COUNTER_DTYPE.itemsize)
    forms = [x for x in prec.split(",") if x]
    tidx = {}
    if forms:
        from pickle_amd import precision

        for form in forms:
            tp = time.perf_counter()
            tidx[form] = precision.traversal_index(idx, base, form)
            log(f"{form} traversal copy + FP32 rows in HBM ({time.perf_counter() - tp:.1f}s)")

    # legs: "exact" (k_search, the reference's heap order), "fast" (k_search_fast),
    # and per precision form "<form>" (fast traversal of the copy + exact FP32 re-rank)
    def handle(leg):
        return idx if leg in ("exact", "fast") else tidx[leg]

    def launch(leg, ef, c=False):
        ctr = d_ctr if c else None
        if leg == "exact":
            idx.search_device(d_q, nq, k, ef, k, d_ids, d_dist, d_cnt, ctr)
        elif leg == "fast":
            idx.search_fast_device(d_q, nq, k, ef, k, d_ids, d_dist, d_cnt, ctr)
        else:
            tidx[leg].search_rerank_fast_device(d_q, nq, k, ef, k, d_ids, d_dist, d_cnt, ctr)

    vbits = {}

    def default_vbits(ef):
        return library_vbits(ef, n)

    def sweep(leg, target, efs):
        h = handle(leg)
        out, chosen = [], None
        for ef in efs:
            if ef < k or (leg != "exact" and ef > 512):
                continue
            while True:
                launch(leg, ef)
                h.sync()
                h.timer_begin()
                launch(leg, ef)
                ms = h.timer_end()
                flags = h.device_error()
                if leg == "exact" and flags == 1 and vbits.get(leg, default_vbits(ef)) < 20:
                    # a query filled its visited table: a larger one per wave
                    vbits[leg] = max(vbits.get(leg, 0), default_vbits(ef)) + 1
                    h.set_tuning(0, vbits[leg], 0)
                    log(f"[{leg}] ef={ef}: visited table full; retrying with 2^{vbits[leg]} slots per wave")
                    continue
                break
            if flags:
                sys.exit(f"[{leg}] ef={ef}: device reported search failure flags {flags} "
                         "(1 visited table full, 2 bad arguments, 4 candidate heap full)")
            ids = d_ids.to_array(np.int32, (nq, k))[:gt_nq]
            r = recall_at_k(ids, gt_ids, k)
            out.append([ef, round(r, 4), round(nq / (ms / 1e3), 1)])
            log(f"[{leg}] ef={ef:4d} recall@{k}={r:.4f} qps={nq / (ms / 1e3):,.0f}")
            if r >= target and chosen is None:
                chosen = ef
                if not args.ef:
                    break
        if chosen is None:
            chosen = out[-1][0]
            log(f"[{leg}] no ef reached recall {target}; using ef={chosen}")
        return chosen, [x[1] for x in out if x[0] == chosen][0], out

    tuned = {}

    def autotune(leg, ef):
        """The library's autotuner at this ef on up to 16,384 of the queries
        (the exact search's tuning cannot change results; the fast mode's
        width can, and the sweep after it measures the recall it gives)."""
        if leg != "exact" and args.fast_tuning:
            w, hb, wpg, bpc = (int(x) for x in args.fast_tuning.split(","))
            for h in [idx, *tidx.values()]:
                h.set_fast_tuning(w, hb, wpg, bpc)
            tuned[leg] = {"width": w, "hash_bits": hb, "waves_per_group": wpg, "blocks_per_cu": bpc, "fixed": True}
            log(f"[{leg}] fixed tuning {tuned[leg]}")
            return
        if args.no_autotune:
            return
        ns = min(nq, 16384)
        if leg == "exact":
            t = idx.autotune(d_q, ns, k, ef)
        else:
            t = idx.autotune_fast(d_q, ns, k, min(ef, 512))
            # the precision forms take the width, filter and workgroup shape;
            # a cap on resident workgroups found for FP32 rows would starve
            # their narrower, more latency-bound gathers
            for h in tidx.values():
                h.set_fast_tuning(t["width"], t["hash_bits"], t["waves_per_group"], 0)
        tuned[leg] = t
        log(f"[{leg}] autotuned at ef={ef}: {t}")

    def timed(leg, ef, steps):
        h = handle(leg)
        for _ in range(args.warmup):
            launch(leg, ef)
        h.sync()
        tw0 = time.perf_counter()
        h.timer_begin()
        for _ in range(steps):
            launch(leg, ef)
        ev_ms = h.timer_end()
        h.sync()
        if h.device_error():
            sys.exit(f"[{leg}] device reported a search failure")
        return time.perf_counter() - tw0, ev_ms / 1e3 / steps

    def counters(leg, ef):
        launch(leg, ef, True)
        handle(leg).sync()
        return d_ctr.to_array(abi.COUNTER_DTYPE, (nq,))

    def alg_bytes_of(ctr, elem, rerank_ef=0):
        """Algorithmic bytes of a launch in the reference Profiler's units: 4 B
        per adjacency entry scanned, 8 B of node metadata per hop, one row per
        distance, the query; the precision forms add the FP32 rows of the ef
        re-ranked candidates and the FP32 query."""
        hops = ctr["hops"].sum(axis=1).astype(np.int64)
        nbr = ctr["neighbors"].sum(axis=1).astype(np.int64)
        dists = ctr["dists"].sum(axis=1).astype(np.int64)
        b = int((4 * nbr + 8 * hops + elem * dim * dists).sum() + nq * elem * dim)
        if rerank_ef:
            b += nq * (rerank_ef * (4 * dim + 4) + 4 * dim)
        return b, hops, nbr, dists

    efs = [args.ef] if args.ef else [int(e) for e in args.efs.split(",")]
    mode = args.mode
    results = {}
    # the exact search: the reference's heap order, the CPU baseline's ef
    ef_x, r_x, sw_x = sweep("exact", args.recall_target, efs)
    autotune("exact", ef_x)  # tuning cannot change the exact search's results
    steps_x = args.steps if mode == "exact" else max(3, args.steps // 4)
    wall_x, ks_x = timed("exact", ef_x, steps_x)
    results["exact"] = {"ef": ef_x, "recall@10": r_x, "qps": round(nq * steps_x / wall_x, 1),
                        "kernel_ms": round(ks_x * 1e3, 4), "sweep_ef_recall_qps": sw_x, "kernel": "k_search"}
    exact_ids = d_ids.to_array(np.int32, (nq, k))
    if mode == "fast":
        autotune("fast", ef_x)
        ef_f, r_f, sw_f = sweep("fast", args.recall_target, efs)
        # a first timing for the precision forms' ratio; the headline is timed last
        steps_f = max(3, args.steps // 4)
        wall_f, ks_f = timed("fast", ef_f, steps_f)
        results["fast"] = {"ef": ef_f, "recall@10": r_f, "qps": round(nq * steps_f / wall_f, 1),
                           "kernel_ms": round(ks_f * 1e3, 4), "sweep_ef_recall_qps": sw_f, "kernel": "k_search_fast"}
    main_leg = mode
    r_main = results[main_leg]["recall@10"]
    for form in forms:
        # the precision form at equal recall: the first ef within 0.002 of the FP32 leg's recall
        ef_p, r_p, sw_p = sweep(form, r_main - 0.002, efs)
        steps_p = max(3, args.steps // 4)
        wall_p, ks_p = timed(form, ef_p, steps_p)
        results[form] = {"ef": ef_p, "recall@10": r_p, "qps": round(nq * steps_p / wall_p, 1),
                         "kernel_ms": round(ks_p * 1e3, 4), "sweep_ef_recall_qps": sw_p,
                         "kernel": "k_quantize + k_search_fast + k_rerank",
                         "speedup_vs_fp32": round(nq * steps_p / wall_p / results[main_leg]["qps"], 3)}
        log(f"{form} traversal + FP32 re-rank: {results[form]['qps']:,.0f} QPS at recall {r_p:.4f} (ef {ef_p}) vs "
            f"FP32 {main_leg} {results[main_leg]['qps']:,.0f} QPS at {r_main:.4f} -> "
            f"{results[form]['speedup_vs_fp32']:.3f}x")

    # the headline: the main leg, timed again last so its launches are the
    # last `steps` of its kernel in a profile
    ef_m = results[main_leg]["ef"]
    ctr_m = counters(main_leg, ef_m)
    wall, kernel_s = timed(main_leg, ef_m, args.steps)
    grid, lds = handle(main_leg).last_launch()  # the timed kernel's launch shape
    # the timed launches' outputs, before any other launch reuses the buffers
    all_ids = d_ids.to_array(np.int32, (nq, k))
    all_d = d_dist.to_array(np.float32, (nq, k))
    qps = nq * args.steps / wall
    results[main_leg]["qps"] = round(qps, 1)
    results[main_leg]["kernel_ms"] = round(kernel_s * 1e3, 4)
    for form in forms:
        results[form]["speedup_vs_fp32"] = round(results[form]["qps"] / qps, 3)
    elem = base.dtype.itemsize
    counted, hops, nbr, dists = alg_bytes_of(ctr_m, elem)
    # algorithmic bytes: the reference algorithm's work at this ef (the exact
    # search's counters); the fast kernel's own counters also include the
    # rows its LDS filter made it read twice
    ctr_ref = counters("exact", ef_m) if main_leg == "fast" else ctr_m
    alg_bytes, hops_r, nbr_r, dists_r = alg_bytes_of(ctr_ref, elem)
    achieved = alg_bytes / kernel_s / 1e9
    khz = abi.wall_clock_khz()
    lat_ms = ctr_m["latency_ticks"].astype(np.float64) / khz
    log(f"timed: {args.steps} x {nq} queries at ef={ef_m} ({main_leg}): {qps:,.0f} QPS, kernel {kernel_s * 1e3:.3f} ms, "
        f"{achieved:.1f} GB/s algorithmic (grid {grid}, LDS {lds} B/workgroup)")
    r_ids = all_ids[:gt_nq]
    if main_leg != "exact":
        # the fast leg against the exact search (the reference's results)
        same = float(np.mean(np.all(all_ids == exact_ids, axis=1)))
        results[main_leg]["ids_identical_to_exact"] = round(same, 4)
        results[main_leg]["recall_minus_exact"] = round(results[main_leg]["recall@10"] - r_x, 4)
        log(f"{main_leg} leg: ids identical to the exact search on {same * 100:.2f}% of queries, recall "
            f"{results[main_leg]['recall@10']:.4f} vs {r_x:.4f}")
    r_d = all_d[:gt_nq].astype(np.float64)
    ok = r_ids >= 0
    rows = base[np.where(ok, r_ids, 0)].astype(np.float64)
    qv = queries[:gt_nq].astype(np.float64)[:, None, :]
    d64 = ((rows - qv) ** 2).sum(-1) if metric == "l2" else -(rows * qv).sum(-1)
    rel = np.abs(r_d - d64) / np.maximum(np.abs(d64), 1e-30)
    dist_err = {"max_rel": float(rel[ok].max()), "mean_rel": float(rel[ok].mean()), "queries": int(gt_nq)}
    log(f"returned distances vs FP64: max relative error {dist_err['max_rel']:.3g}, "
        f"mean {dist_err['mean_rel']:.3g}")

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
    # graph and data, at the exact search's ef (the reference's own operating
    # point at the recall target)
    cpu = None
    hostcpu = cpu_info()
    if not args.no_cpu_search:
        cnq = min(nq, args.cpu_queries) if args.cpu_queries else nq
        if cnq < nq:
            queries[:cnq].tofile(q_path + ".sample")
            cq_path = q_path + ".sample"
        else:
            cq_path = q_path
        out = os.path.join(work, "cpu_res")
        st = cpu_search(binary, graph_path, base_path, cq_path, n, cnq, dim, dtype_name, k, ef_x, threads,
                        args.cpu_repeat, out)
        cpu_ids = read_ref_results(f"{out}.ef{ef_x}", cnq, k)
        r_n = min(cnq, gt_nq)
        cpu_recall = recall_at_k(cpu_ids[:r_n], gt_ids[:r_n], k)
        same = float(np.mean(np.all(cpu_ids == exact_ids[:cnq], axis=1)))
        cpu = {"value": round(st["qps"], 1), "unit": "queries/s", "cores": threads, "kind": "reference",
               "sample": f"{cnq} of the {nq} queries, ef={ef_x}, k={k}, best of {args.cpu_repeat} runs, reference "
                         f"AnnSearch compiled from source ({isa} distance path), OpenMP on {threads} threads "
                         f"(every core this process may use) of {hostcpu['model']}, "
                         f"{'reference-built' if build == 'cpu' else 'GPU-built'} graph",
               "qps_per_thread": round(st["qps"] / threads, 1), "recall@10": round(cpu_recall, 4),
               "ids_identical_to_gpu_exact": round(same, 4), "host": hostcpu,
               "graph_build_seconds": round(build_info["build_seconds"], 2),
               "graph_built_by": "the reference on the host cores (build_threads)" if build_info.get("where", "cpu")
               == "cpu" else "the GPU", "build_threads": build_info["threads"]}
        log(f"CPU reference: {st['qps']:,.0f} QPS on {threads} threads at ef={ef_x}, recall {cpu_recall:.4f}, "
            f"GPU exact ids identical on {same * 100:.2f}% of queries; GPU {main_leg} = "
            f"{qps / st['qps']:.1f}x")

    traffic, traffic_src, l2_hit = profiled_traffic(args.workload, n, nq, k, ef_m, main_leg)
    cfg = {"workload": wl["desc"] if (n, nq) == (wl["n"], wl["nq"]) else f"{wl['desc']} (n={n}, nq={nq})",
           "n": n, "nq": nq, "dim": dim, "M": M, "efC": efc, "k": k, "ef": ef_m, "recall@10": r_main,
           "recall_queries": gt_nq, "graph": "reference CPU build" if build == "cpu" else
           f"GPU batched build ({build_info['build_seconds']:.1f}s, efC {build_info.get('efc', efc)}, "
           + ("rows capped at D" if build_info.get('capped_rows') else
              f"reference rows up to {build_info.get('row_limit', 511)} ids") +
           (", insert batches replayed as HIP graphs" if build_info.get("hip_graphs") else "") + ")",
           "search": {"exact": "exact (reference heap order)",
                      "fast": "fast mode (sorted register beam, width-w expansion, LDS visited filter)"}[main_leg],
           "p50_query_latency_ms": round(float(np.percentile(lat_ms, 50)), 4),
           "p99_query_latency_ms": round(float(np.percentile(lat_ms, 99)), 4),
           "distance_error_vs_fp64": dist_err,
           "batch_latency_ms": round(wall / args.steps * 1e3, 4),
           "launch": {"kernel": results[main_leg]["kernel"], "grid": grid, "lds_bytes_per_workgroup": lds},
           "legs": results}
    if tuned:
        cfg["autotuned"] = tuned
    line = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype_name, "data": "synthetic", "config": cfg,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src, "l2_hit_rate": l2_hit,
                     "kernel": results[main_leg]["kernel"], "kernel_ms": round(kernel_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "algorithmic_per_query": {"hops": round(float(hops_r.mean()), 2),
                                               "neighbors": round(float(nbr_r.mean()), 2),
                                               "dists": round(float(dists_r.mean()), 2)},
                     "kernel_counted_bytes_per_launch": counted,
                     "kernel_per_query": {"hops": round(float(hops.mean()), 2),
                                          "neighbors": round(float(nbr.mean()), 2),
                                          "dists": round(float(dists.mean()), 2)}},
        "cpu_baseline": cpu,
        "ground_truth": ground_truth,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
