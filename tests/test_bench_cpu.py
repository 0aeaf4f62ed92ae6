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
