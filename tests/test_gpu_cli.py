"""End-to-end drop-in run of the BUSseq_gpu executable on the GPU: the
reference's command line, progress lines and output files (README.md:17-49)."""
import os
import subprocess

import numpy as np
import pytest

import busseq_gpu_amd as bq

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "busseq_gpu_amd", "bin", "BUSseq_gpu")

OUTS = ["_post_Z.txt", "_post_X.txt", "_post_gamma.txt", "_post_alpha.txt", "_post_L.txt", "_post_beta.txt",
        "_post_p.txt", "_post_nu.txt", "_post_delta.txt", "_post_phi.txt", "_post_W.txt", "_post_pi.txt",
        "_BIC.txt"]
PRES = ["_preserved_%s.txt" % n for n in ("gamma", "alpha", "L", "beta", "p", "nu", "delta", "phi", "W", "pi")]


def test_cli_end_to_end_matches_library(tmp_path):
    n_b = [20, 15, 15]
    G, K = 60, 3
    Y = bq.simulate(n_b, G, K, 21)
    (tmp_path / "dim.txt").write_text(" ".join(map(str, n_b)) + "\n")
    (tmp_path / "count.txt").write_text("\n".join(" ".join(map(str, r)) for r in Y) + "\n")
    print("[cli] running BUSseq_gpu", flush=True)
    p = subprocess.run([EXE, "-B", "3", "-N", str(tmp_path / "dim.txt"), "-G", str(G), "-K", str(K), "-s", "13579",
                        "-c", str(tmp_path / "count.txt"), "-i", "30", "-b", "15", "-u", "5", "-p",
                        "-o", str(tmp_path / "cli")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [l for l in p.stdout.splitlines() if l]
    assert lines[:4] == ["Seed is 13579", "Start reading file.", "Start initialization.", "Start MCMC."]
    assert "Burnin completed. Start recording iterations." in lines
    assert "Start posterior inference." in lines and "Start writing output files." in lines
    r = bq.run(Y, n_b, K, seed=13579, n_iter=30, n_burnin=15, n_unchanged=5, keep_traces=True,
               prefix=str(tmp_path / "lib"))
    for name in OUTS + PRES:
        a = (tmp_path / ("cli" + name)).read_text()
        b = (tmp_path / ("lib" + name)).read_text()
        assert a == b, name
    X = np.loadtxt(tmp_path / "cli_post_X.txt", dtype=np.int64)
    Z = np.loadtxt(tmp_path / "cli_post_Z.txt", dtype=np.int64)
    assert X.shape == (50, 60) and np.all(X[Y > 0] == Y[Y > 0]) and np.all(Z[Y > 0] == 0)
    assert np.all((X[(Y == 0) & (Z == 0)]) == 0)
    W = np.loadtxt(tmp_path / "cli_post_W.txt", dtype=np.int64)
    np.testing.assert_array_equal(W, r["W"])
    assert np.loadtxt(tmp_path / "cli_preserved_delta.txt").shape == (50, 15)
