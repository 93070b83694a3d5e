"""GPU parity on the benchmark workloads, at their full shapes, against the CPU
oracle (busseq_gpu_amd/workloads.py builds the same data sets bench.py times).

  c2  N=20000 (8000/6000/4000/2000), G=3000, K=7: three sweeps, every state
      bit-identical to the oracle's (integer draws exact, parameters equal).
  c4  the K = 3..10 sweep with 4 chains each on the c1 data, 32 chains in one
      batched sampler: every chain against its own oracle for two sweeps, and
      the posterior summaries and BIC of one chain per K equal to a plain run.
  c5  the 16 heterogeneous data sets in one batched sampler: every chain
      against its own oracle for two sweeps.
  c3  N=150000, G=5000, K=12, B=8: one sweep against the oracle, plus the
      state's invariants (observed counts never imputed, W in [0, K), pi rows
      summing to one, phi > 0, gamma1 < 0) and a bit-identical rerun.

Every workload runs in the phase the benchmark times (past burn-in and past
n_unchanged: p/tau0 updates and posterior accumulation on every sweep) with
the tile size the autotuner picks for it on the MI355X.
"""
import numpy as np
import pytest

import busseq_gpu_amd as bq
from busseq_gpu_amd import workloads as wl
from oracle import pyoracle as po
from parity import compare, oracle_threads

pytestmark = pytest.mark.gpu

PHASE = dict(n_iter=4000, n_burnin=0, n_unchanged=0)


def oracle_for(p, chain, tile_rows, precision=32):
    return po.Oracle(p["Y"], p["n_b"], p["K"], seed=p.get("seed", wl.SEED), chain=chain, precision=precision,
                     tile_rows=tile_rows, threads=oracle_threads(), **PHASE)


def test_c2_full_shape_three_sweeps_match_oracle():
    cfg = wl.CONFIGS["c2"]
    Y = wl.counts(bq, "c2")
    tr = 32
    print(f"[c2] N={Y.shape[0]} G={Y.shape[1]} K={cfg['K']} TR={tr}", flush=True)
    o = oracle_for(dict(Y=Y, n_b=cfg["n_b"], K=cfg["K"]), 0, tr)
    with bq.Sampler(Y, cfg["n_b"], cfg["K"], seed=wl.SEED, tile_rows=tr, **PHASE) as s:
        compare(s.state(), o.state(), "c2 init")
        for step in (1, 2):
            s.sweeps(step)
            o.sweeps(step)
            print(f"[c2] iteration {o.iteration}", flush=True)
            compare(s.state(), o.state(), f"c2 iteration {o.iteration}")
        r = s.finish()
    p = o.posterior()
    for k in ("gamma", "alpha", "L", "beta", "ptau", "nu", "delta", "phi", "pi"):
        np.testing.assert_allclose(r[k], p[k], rtol=1e-12, atol=1e-300, err_msg=f"c2 posterior {k}")
    np.testing.assert_array_equal(r["W"], p["W"])


def test_c4_k_sweep_every_chain_matches_oracle_and_plain_runs():
    tr = 64
    probs = wl.c4_problems(bq, **PHASE)
    n_chains = sum(p["chains"] for p in probs)
    print(f"[c4] {len(probs)} K values x {wl.C4_CHAINS} chains = {n_chains} chains, TR={tr}", flush=True)
    sb = bq.Sampler.batch(probs, tile_rows=tr)
    assert sb.chains == n_chains == 32
    sb.sweeps(2)
    j = 0
    for p in probs:
        for c in range(p["chains"]):
            o = oracle_for(p, c, tr)
            o.sweeps(2)
            compare(sb.state(j), o.state(), f"c4 K={p['K']} chain {c}")
            o.close()
            j += 1
    # the posterior summaries and BIC of chain 1 of each K equal a plain run of that chain
    j = 0
    for p in probs:
        rb = sb.finish(j + 1)
        with bq.Sampler(p["Y"], p["n_b"], p["K"], seed=p["seed"], chain0=1, tile_rows=tr, **PHASE) as sp:
            sp.sweeps(2)
            rp = sp.finish()
        for k in ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi", "ptau", "L", "W"):
            assert np.array_equal(rb[k], rp[k]), f"c4 K={p['K']} posterior {k}"
        assert rb["bic"] == rp["bic"] and np.isfinite(rb["bic"]), f"c4 K={p['K']} BIC"
        print(f"[c4] K={p['K']} BIC {rb['bic']:.1f}", flush=True)
        j += p["chains"]
    sb.close()


def test_c5_sixteen_data_sets_every_chain_matches_oracle():
    tr = 128
    probs = wl.c5_problems(bq, **PHASE)
    print(f"[c5] {len(probs)} data sets, {sum(p['Y'].size for p in probs)} cells x genes, TR={tr}", flush=True)
    with bq.Sampler.batch(probs, tile_rows=tr) as sb:
        sb.sweeps(2)
        for j, p in enumerate(probs):
            o = oracle_for(p, 0, tr)
            o.sweeps(2)
            compare(sb.state(j), o.state(), f"c5 data set {j} (N={p['Y'].shape[0]} G={p['Y'].shape[1]} "
                                            f"B={len(p['n_b'])} K={p['K']})")
            o.close()


@pytest.mark.timeout(1200)
def test_c3_atlas_shape_one_sweep_matches_oracle():
    cfg = wl.CONFIGS["c3"]
    Y = wl.counts(bq, "c3")
    K, tr = cfg["K"], 128
    N, G = Y.shape
    print(f"[c3] N={N} G={G} K={K} B={len(cfg['n_b'])} TR={tr} zeros={np.mean(Y == 0):.3f}", flush=True)
    with bq.Sampler(Y, cfg["n_b"], K, seed=wl.SEED, tile_rows=tr, **PHASE) as s:
        s.sweeps(1)
        gs = s.state()
    # invariants of the state at full size
    obs = Y > 0
    assert np.array_equal(gs["X"][obs], Y[obs]) and not gs["Z"][obs].any()
    assert not gs["X"][~obs][gs["Z"][~obs] == 0].any()
    assert gs["W"].min() >= 0 and gs["W"].max() < K
    assert np.allclose(gs["pi"].sum(1), 1.0, atol=1e-12) and np.all(gs["pi"] > 0)
    assert np.all(gs["phi"] > 0) and np.all(gs["gamma"][1] < 0)
    # a rerun gives the same bits
    with bq.Sampler(Y, cfg["n_b"], K, seed=wl.SEED, tile_rows=tr, **PHASE) as s2:
        s2.sweeps(1)
        compare(s2.state(), gs, "c3 rerun")
    print("[c3] GPU sweep done; oracle sweep", flush=True)
    o = oracle_for(dict(Y=Y, n_b=cfg["n_b"], K=K), 0, tr)
    del Y
    o.sweeps(1)
    compare(gs, o.state(), "c3 iteration 1")
