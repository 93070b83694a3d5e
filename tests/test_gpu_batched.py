"""Batched form: several problems (different count matrices, K, B, seeds) and
chains advanced in the same launches.  Each chain must reproduce the plain run
with the same Philox key bit for bit (integer draws and parameters), and its
posterior summaries and BIC must equal the plain run's."""
import numpy as np
import pytest

import busseq_gpu_amd as bq

pytestmark = pytest.mark.gpu

KEYS_I = ("X", "Z", "W", "L")
KEYS_F = ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi", "ptau")


def same(a, b, where):
    for k in KEYS_I:
        assert np.array_equal(a[k], b[k]), f"{where}: {k}"
    for k in KEYS_F:
        assert np.array_equal(a[k], b[k]), f"{where}: {k}"


def problems():
    out = []
    specs = [([20, 15], 80, 3, 11), ([12, 18, 10], 64, 4, 12), ([30, 25, 20, 15], 120, 5, 13), ([9, 9], 40, 2, 14)]
    for n_b, G, K, ds in specs:
        Y = bq.simulate(n_b, G, K, ds)
        out.append(dict(Y=Y, n_b=n_b, K=K, seed=1000 + ds, chains=2, chain0=3, n_iter=30, n_burnin=15, n_unchanged=5))
    return out


@pytest.mark.parametrize("prec", [32, 64])
def test_heterogeneous_batch_equals_plain_runs(prec):
    probs = problems()
    print(f"[batched] {len(probs)} problems x 2 chains, prec={prec}", flush=True)
    sb = bq.Sampler.batch(probs, precision=prec)
    sb.sweeps(30)
    j = 0
    for p in probs:
        for c in range(p["chains"]):
            kw = {k: v for k, v in p.items() if k not in ("Y", "n_b", "K", "chains", "chain0")}
            with bq.Sampler(p["Y"], p["n_b"], p["K"], chains=1, chain0=p["chain0"] + c, precision=prec, **kw) as sp:
                sp.sweeps(30)
                same(sb.state(j), sp.state(), f"chain {j}")
                rb, rp = sb.finish(j), sp.finish()
                for k in ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi", "ptau", "L", "W"):
                    assert np.array_equal(rb[k], rp[k]), f"chain {j} posterior {k}"
                assert rb["bic"] == rp["bic"]
            j += 1
    sb.close()


def test_batch_of_wide_and_narrow_k_equals_plain_runs():
    # chains with K > 16 (grouped subtype draw) beside chains with K <= 16 in one batch
    specs = [([26, 24], 90, 20, 21), ([15, 15], 60, 4, 22), ([40, 38], 70, 37, 23)]
    probs = []
    for n_b, G, K, ds in specs:
        Y = bq.simulate(n_b, G, K, ds)
        probs.append(dict(Y=Y, n_b=n_b, K=K, seed=2000 + ds, chains=2, n_iter=20, n_burnin=8, n_unchanged=3))
    print(f"[batched] K = {[p['K'] for p in probs]} x 2 chains", flush=True)
    sb = bq.Sampler.batch(probs)
    sb.sweeps(12)
    j = 0
    for p in probs:
        for c in range(p["chains"]):
            kw = {k: v for k, v in p.items() if k not in ("Y", "n_b", "K", "chains")}
            with bq.Sampler(p["Y"], p["n_b"], p["K"], chains=1, chain0=c, **kw) as sp:
                sp.sweeps(12)
                same(sb.state(j), sp.state(), f"chain {j} (K={p['K']})")
            j += 1
    sb.close()


def test_k_sweep_bic_per_k():
    n_b = [40, 30, 30]
    Y = bq.simulate(n_b, 100, 4, 77)
    Ks = (2, 3, 4, 5, 6)
    probs = [dict(Y=Y, n_b=n_b, K=K, seed=5, chains=2, n_iter=40, n_burnin=20, n_unchanged=5) for K in Ks]
    sb = bq.Sampler.batch(probs)
    sb.sweeps(40)
    bic = [[sb.finish(2 * q + c)["bic"] for c in range(2)] for q in range(len(probs))]
    sb.close()
    assert all(np.isfinite(b).all() for b in bic)
    for q in (0, len(Ks) - 1):
        with bq.Sampler(Y, n_b, Ks[q], seed=5, chain0=1, n_iter=40, n_burnin=20, n_unchanged=5) as sp:
            sp.sweeps(40)
            assert sp.finish()["bic"] == bic[q][1]


def test_autotune_returns_a_candidate_tile_size_that_the_oracle_reproduces():
    from oracle import pyoracle as po
    n_b, G, K = [40, 30, 30], 150, 3
    Y = bq.simulate(n_b, G, K, 21)
    tr, sps = bq.autotune(Y, n_b, K, seed=9, sweeps=5)
    assert tr in (16, 32, 64, 128) and sps > 0
    tr_b, _ = bq.autotune_batch(problems(), sweeps=3)
    assert tr_b in (16, 32, 64, 128)
    # a chain run with the chosen tile size equals the oracle with the same grouping
    with bq.Sampler(Y, n_b, K, seed=9, tile_rows=tr, n_iter=20, n_burnin=10, n_unchanged=3) as s:
        s.sweeps(6)
        g = s.state()
    o = po.Oracle(Y, n_b, K, seed=9, n_iter=20, n_burnin=10, n_unchanged=3, tile_rows=tr)
    o.sweeps(6)
    same(g, o.state(), f"tile rows {tr}")
