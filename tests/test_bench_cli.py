"""bench.py's command line and workload builders, on the CPU (no GPU call is made:
every case below exits or returns before a sampler is created)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_batch_only_needs_a_fixed_tile_size():
    r = run_bench("--batch-only", "c4")
    assert r.returncode == 2
    assert "--tile-rows" in r.stderr and r.stdout == ""


def test_batched_workloads_have_their_stated_shapes():
    import bench
    import busseq_gpu_amd as bq
    from busseq_gpu_amd.workloads import c5_datasets

    p4 = bench.c4_problems(bq)
    assert [p["K"] for p in p4] == list(range(3, 11))
    assert all(p["chains"] == 4 and p["Y"].shape == (1000, 3000) for p in p4)
    assert all(p["Y"] is p4[0]["Y"] for p in p4)           # one data set, eight values of K
    assert all(p["n_iter"] == 4000 and p["n_burnin"] == 0 and p["n_unchanged"] == 0 for p in p4)

    p5 = bench.c5_problems(bq)
    sets = c5_datasets()
    assert len(p5) == 16
    for p, d in zip(p5, sets):
        assert p["Y"].shape == (sum(d["n_b"]), d["G"]) and p["K"] == d["K"] and list(p["n_b"]) == list(d["n_b"])
        assert p["Y"].dtype == np.int32 and (p["Y"] >= 0).all()
    assert sum(int(p["Y"].size) for p in p5) == 375444440
