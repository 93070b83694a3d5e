"""Posterior-level checks on the GPU.

Traces: every recorded iteration (the -p output of BUSseq_gpu.cu:892-936,
1846-2018) averages to the posterior mean the sampler accumulates and ends in
the chain's final state.

Mixed precision against FP64 (the agreement the mixed-precision form must
keep): on the c1 shape (B=4, N=1000, G=3000, K=5) four FP32 chains run in one
batched sampler; the chain of lowest BIC and an FP64 chain with the same
Philox key recover the true subtypes (ARI >= 0.99, also against each other),
and their posterior means agree within 3 Monte-Carlo standard errors for all
but a few percent of the parameters (standard errors from the effective
sample size, Geyer's initial positive sequence).  The cell effects delta of a
batch and the batch effects nu of that batch have one slowly mixing direction
-- all delta of the batch up by c, all its nu down by c, which moves only the
batch's first cell (its delta is fixed at 0, cu:766-772); likewise all alpha
up and all delta down moves only the first cells -- so alpha, delta and nu are
compared after removing their mean (over genes for alpha, over the batch for
delta and nu) in every recorded iteration, and those means are bounded
separately.  The dropout curve gamma
(2 parameters per batch) mixes slowly with the imputed counts -- a few
effective draws per chain, too few for a Monte-Carlo error -- and is held to
3 posterior standard deviations instead.
"""
import numpy as np
import pytest

import busseq_gpu_amd as bq
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_traces_match_posterior_and_state():
    n_b, G, K = [30, 25, 20], 200, 3
    Y = bq.simulate(n_b, G, K, 77)
    with bq.Sampler(Y, n_b, K, seed=5, n_iter=40, n_burnin=15, n_unchanged=5, keep_traces=True) as s:
        s.sweeps(40)
        st = s.state()
        post = s.finish()
    tr = post["traces"]
    assert post["n_pres"] == 25 and tr["alpha"].shape == (G, 25)
    for key in ("alpha", "delta", "pi", "gamma"):
        np.testing.assert_allclose(tr[key].mean(1), np.ravel(post[key]), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(tr["beta"].mean(1), post["beta"][1:].ravel(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(tr["nu"].mean(1), post["nu"][1:].ravel(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(tr["phi"].mean(1), post["phi"].ravel(), rtol=1e-12, atol=1e-12)
    assert np.array_equal(tr["alpha"][:, -1], st["alpha"])
    assert np.array_equal(tr["delta"][:, -1], st["delta"])
    assert np.array_equal(tr["W"][:, -1], st["W"])
    assert np.array_equal(tr["beta"][:, -1], st["beta"][1:].ravel())
    assert np.array_equal(tr["L"][:, -1], st["L"][1:].ravel())


def test_traces_of_a_partial_run_match_the_oracle():
    # finish after fewer sweeps than n_iter: the result holds the recorded sweeps
    # only, with the stride it reports (n_pres), and they equal the oracle's traces
    # (cu:892-936) column by column
    n_b, G, K = [30, 25, 20], 200, 3
    Y = bq.simulate(n_b, G, K, 78)
    kw = dict(seed=6, n_iter=40, n_burnin=15, n_unchanged=5)
    o = po.Oracle(Y, n_b, K, keep_traces=True, **kw)
    o.sweeps(27)
    ot, op = o.traces(), o.posterior()
    with bq.Sampler(Y, n_b, K, keep_traces=True, **kw) as s:
        s.sweeps(27)
        post = s.finish()
    nrec = 27 - 15
    assert post["n_pres"] == nrec
    for key in po.TRACE_KEYS:
        assert post["traces"][key].shape == (ot[key].shape[0], nrec), key
        assert np.array_equal(post["traces"][key], ot[key][:, :nrec]), key
    for key in ("alpha", "delta", "phi", "pi", "gamma", "beta", "nu", "L", "ptau"):
        np.testing.assert_allclose(post[key], op[key], rtol=1e-12, atol=1e-300, err_msg=key)
    np.testing.assert_allclose(post["traces"]["alpha"].mean(1), post["alpha"], rtol=1e-12)


def test_streamed_traces_over_many_blocks_match_the_oracle():
    # traces stream to host memory in blocks of 32 sweeps through a device ring of
    # two blocks: runs that wrap the ring several times, two problems in one batch
    # with different burn-in and run lengths (their blocks are recorded from
    # different sweeps), and sweeps launched one kernel at a time (the per-kernel
    # timing path) in the middle; every chain's traces equal its oracle's
    specs = [([24, 20], 90, 3, 31, 150, 10), ([18, 16, 14], 70, 4, 32, 120, 53)]
    probs, oracles = [], []
    for n_b, G, K, ds, n_iter, n_burnin in specs:
        Y = bq.simulate(n_b, G, K, ds)
        kw = dict(seed=ds, n_iter=n_iter, n_burnin=n_burnin, n_unchanged=5)
        probs.append(dict(Y=Y, n_b=n_b, K=K, keep_traces=True, **kw))
        o = po.Oracle(Y, n_b, K, keep_traces=True, **kw)
        o.sweeps(150)
        oracles.append(o)
    with bq.Sampler.batch(probs) as s:
        s.sweeps(60)
        s.kernel_times(7)        # sweeps 60..66: across the boundary at 64
        s.sweeps(83)
        assert s.iteration == 150
        for j, (o, p) in enumerate(zip(oracles, probs)):
            post = s.finish(j)
            P = p["n_iter"] - p["n_burnin"]
            assert post["n_pres"] == P
            ot = o.traces()
            for key in po.TRACE_KEYS:
                assert np.array_equal(post["traces"][key], ot[key]), (j, key)


def mc_se(tr, ess=False):
    """Monte-Carlo standard error of each row's mean, tau from Geyer's initial positive
    sequence; with ess=True also the effective sample size n / tau"""
    x = tr - tr.mean(1, keepdims=True)
    n = x.shape[1]
    f = np.fft.rfft(x, 2 * n, axis=1)
    acov = np.fft.irfft(f * np.conj(f), axis=1)[:, :n] / n
    var = acov[:, 0]
    rho = acov / np.maximum(var[:, None], 1e-300)
    m = (n - 1) // 2
    pairs = rho[:, 0:2 * m:2] + rho[:, 1:2 * m:2]
    keep = np.cumprod(pairs > 0, axis=1)
    tau = np.maximum(-1.0 + 2.0 * (pairs * keep).sum(1), 1.0)
    se = np.sqrt(var * tau / n)
    return (se, n / tau) if ess else se


def test_mixed_precision_posterior_agrees_with_fp64():
    from sklearn.metrics import adjusted_rand_score as ari
    n_b, G, K = [300, 300, 200, 200], 3000, 5
    Y, W_true, _ = bq.simulate(n_b, G, K, 1001, want_truth=True)
    run = dict(seed=13579, n_iter=4000, n_burnin=2000, n_unchanged=500, keep_traces=True)
    with bq.Sampler(Y, n_b, K, chains=4, precision=32, **run) as s:
        s.sweeps(4000)
        posts = [s.finish(c) for c in range(4)]
    best = int(np.argmin([p["bic"] for p in posts]))
    p32 = posts[best]
    with bq.Sampler(Y, n_b, K, chain0=best, precision=64, **run) as s:
        s.sweeps(4000)
        p64 = s.finish()
    print(f"\nchain {best}: ARI(truth) fp32 {ari(W_true, p32['W']):.4f} fp64 {ari(W_true, p64['W']):.4f} "
          f"fp32-fp64 {ari(p32['W'], p64['W']):.4f}; BIC {p32['bic']:.1f} / {p64['bic']:.1f}")
    assert ari(W_true, p32["W"]) >= 0.99 and ari(W_true, p64["W"]) >= 0.99
    assert ari(p32["W"], p64["W"]) >= 0.99
    batch = np.repeat(np.arange(len(n_b)), n_b)

    def centred(p, key):
        tr = p["traces"][key].copy()
        if key == "alpha":        # minus the mean over genes, per iteration
            tr -= tr.mean(0)
        if key == "delta":        # minus the batch mean of the cells' delta, per iteration
            for b in range(len(n_b)):
                tr[batch == b] -= tr[batch == b].mean(0)
        if key == "nu":           # rows (b-1)*G + g for b >= 1: minus the batch mean over genes
            tr = tr.reshape(len(n_b) - 1, G, -1)
            tr = (tr - tr.mean(1, keepdims=True)).reshape((len(n_b) - 1) * G, -1)
        return tr

    for key in ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi"):
        t32, t64 = centred(p32, key), centred(p64, key)
        (s32, e32), (s64, e64) = mc_se(t32, ess=True), mc_se(t64, ess=True)
        se = np.sqrt(s32 ** 2 + s64 ** 2)
        sd = np.sqrt(0.5 * (t32.var(1) + t64.var(1)))
        d = np.abs(t32.mean(1) - t64.mean(1))
        live = se > 0
        frac = np.mean(d[live] <= 3 * se[live])
        print(f"{key:6s} {live.sum():6d} parameters: {frac:.4f} within 3 MC s.e. (median ESS "
              f"{np.median(np.minimum(e32, e64)[live]):.0f}); largest difference "
              f"{np.max(d[live] / sd[live]):.2f} posterior sd")
        if live.sum() >= 100:
            assert frac >= 0.97, key
        else:
            # a handful of slowly mixing parameters (the dropout curve of each batch,
            # a few effective draws per chain): agreement within 3 posterior sd
            assert np.all(d[live] <= 3 * sd[live]), key
        assert np.all(d[~live] == 0), key
    shift = [abs(p32["delta"][batch == b].mean() - p64["delta"][batch == b].mean()) for b in range(len(n_b))]
    shift.append(abs(p32["alpha"].mean() - p64["alpha"].mean()))
    print("shift of mean delta per batch and of mean alpha:", np.round(shift, 4))
    assert max(shift) < 0.05
