"""bench.py's own logic on the CPU: the recall definition, the host ground
truth, the profiled-traffic lookup and the one-GPU contract."""
import json
import os
import subprocess
import sys

import numpy as np

import bench
from oracle import oracle as O
from pickle_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_recall_counts_label_matches_only():
    # benchmark/hnsw_pickle.cpp:94-112: an id counts if it is among the first k labels
    gt = np.array([[1, 2, 3, 4], [5, 6, 7, 8]])
    ids = np.array([[3, 2, 9], [8, 5, 6]])
    assert bench.recall_at_k(ids, gt, 3) == (2 + 2) / 6


def test_host_ground_truth_equals_the_oracle():
    for metric, om in (("l2", O.L2), ("ip", O.IP)):
        if metric == "l2":
            b, q = datasets.gaussian_mixture(3000, 40, 24, 8, 7, kind="f32")
        else:
            b, q = datasets.vmf_mixture(3000, 40, 24, 8, 7)
        want, _ = O.brute_force(b, q, 15, om)
        np.testing.assert_array_equal(bench.host_ground_truth(b, q, 15, metric), want)


def test_profiled_traffic_lookup(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "PROFILES", str(tmp_path))
    rec = {"workload": "w", "n": 10, "nq": 2, "k": 10, "ef": 96, "traffic_bytes": 1234}
    (tmp_path / "k_search_w_ef96_pmc.md").write_text("text\ntraffic-json: " + json.dumps(rec) + "\n")
    assert bench.profiled_traffic("w", 10, 2, 10, 96)[0] == 1234
    assert bench.profiled_traffic("w", 11, 2, 10, 96) == (None, None, None)  # another size
    assert bench.profiled_traffic("w", 10, 2, 10, 64) == (None, None, None)  # another ef


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def _counter_csv(path, rows):
    path.mkdir(parents=True, exist_ok=True)
    lines = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
    lines += [f'{d},"{k}",{c},{v}' for d, k, c, v in rows]  # quoted, as rocprofv3 writes names with commas
    (path / "run_counter_collection.csv").write_text("\n".join(lines) + "\n")


def test_pmc_summary_calibrates_traffic_and_reads_l2_hits(tmp_path):
    """--summarize-pmc: bytes scaled by what the counters read for the 256 MiB
    copy, only the timed dispatches of k_search, and the L2 hit rate of the
    TCC pass, recorded in the traffic line the bench reads back."""
    kib = bench.CALIB_BYTES / 1024
    search = "void pk::k_search<pk::IntDist<0, 1, 1>, true>(pk::SearchArgs)"
    copy = "pk::k_copy16(void*)"
    # FETCH reads half the copy's bytes; the first k_search dispatch is an untimed one
    _counter_csv(tmp_path / "f", [(1, search, "FETCH_SIZE", 999999), (2, search, "FETCH_SIZE", 1000),
                                  (3, copy, "FETCH_SIZE", kib / 2)])
    _counter_csv(tmp_path / "w", [(1, search, "WRITE_SIZE", 999999), (2, search, "WRITE_SIZE", 300),
                                  (3, copy, "WRITE_SIZE", kib)])
    _counter_csv(tmp_path / "l", [(1, search, "TCC_HIT_sum", 1), (1, search, "TCC_MISS_sum", 1),
                                  (2, search, "TCC_HIT_sum", 25), (2, search, "TCC_MISS_sum", 75)])
    bl = {"steps": 1, "config": {"workload": "w", "ef": 96, "k": 10, "nq": 2, "n": 10, "recall@10": 0.95},
          "roofline": {"algorithmic_bytes_per_launch": 1024000}}
    (tmp_path / "b.json").write_text(json.dumps(bl) + "\n")
    out = tmp_path / "p.md"
    args = type("A", (), dict(bench_json=str(tmp_path / "b.json"), cmd=["cmd"], summarize_trace=None,
                              summarize_pmc=(str(tmp_path / "f"), str(tmp_path / "w")),
                              summarize_l2=str(tmp_path / "l"), workload="w", out=str(out)))
    bench.summarize(args)
    rec = json.loads([x for x in out.read_text().splitlines() if x.startswith("traffic-json: ")][0][14:])
    assert rec["fetch_bytes"] == 1000 * 1024 * 2 and rec["write_bytes"] == 300 * 1024
    assert rec["traffic_bytes"] == rec["fetch_bytes"] + rec["write_bytes"]
    assert rec["l2_hit_rate"] == 0.25


def test_pmc_summary_keeps_kernels_apart(tmp_path):
    """A fast-mode line reads only k_search_fast's dispatches, not the exact
    k_search's of the same process (their short names share a prefix)."""
    kib = bench.CALIB_BYTES / 1024
    exact = "void pk::k_search<pk::FloatDist<2, 1, 8>, false>(pk::SearchArgs)"
    fast = "void pk::k_search_fast<pk::FloatDist<2, 1, 8>, false, 8>(pk::FastArgs)"
    copy = "pk::k_copy16(void*)"
    _counter_csv(tmp_path / "f", [(1, exact, "FETCH_SIZE", 5000), (2, fast, "FETCH_SIZE", 700),
                                  (3, exact, "FETCH_SIZE", 5000), (4, fast, "FETCH_SIZE", 900),
                                  (5, copy, "FETCH_SIZE", kib / 2)])
    _counter_csv(tmp_path / "w", [(1, exact, "WRITE_SIZE", 50), (2, fast, "WRITE_SIZE", 1),
                                  (3, exact, "WRITE_SIZE", 50), (4, fast, "WRITE_SIZE", 3),
                                  (5, copy, "WRITE_SIZE", kib)])
    bl = {"steps": 2, "config": {"workload": "w", "ef": 320, "k": 10, "nq": 2, "n": 10, "recall@10": 0.95},
          "roofline": {"algorithmic_bytes_per_launch": 1024000, "kernel": "k_search_fast"}}
    (tmp_path / "b.json").write_text(json.dumps(bl) + "\n")
    out = tmp_path / "p.md"
    args = type("A", (), dict(bench_json=str(tmp_path / "b.json"), cmd=["cmd"], summarize_trace=None,
                              summarize_pmc=(str(tmp_path / "f"), str(tmp_path / "w")),
                              summarize_l2=None, workload="w", out=str(out)))
    bench.summarize(args)
    rec = json.loads([x for x in out.read_text().splitlines() if x.startswith("traffic-json: ")][0][14:])
    assert rec["kernel"] == "k_search_fast"
    assert rec["fetch_bytes"] == 800 * 1024 * 2 and rec["write_bytes"] == 2 * 1024
