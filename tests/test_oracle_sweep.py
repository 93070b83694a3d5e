"""One oracle sweep recomputed independently with scipy.

From the state before a sweep, the state after it and the proposals the
oracle drew, every Metropolis-Hastings sum (alpha, beta, nu, delta, phi,
gamma0, gamma1; BUSseq_gpu.cu:524-818) and the subtype logits are recomputed
as differences of scipy log-pmfs, and every accept/reject decision is replayed
with the uniform of its Philox stream.  Plus determinism across thread
counts and recovery of synthetic truth.
"""
import numpy as np
import pytest
import scipy.special as sc
import scipy.stats as st

import busseq_gpu_amd as bq
from oracle import pyoracle as po

PH = dict(W=8, ALPHA_A=12, GAMMA0_A=14, GAMMA1_A=16, L=17, BETA_A=20, NU_A=22, DELTA_A=24, PHI_A=26)
SEED = 4242


def lnb(x, mu, phi):
    return st.nbinom.logpmf(x, phi, phi / (phi + mu))


def u(a, b, t, phase):
    return po.stream("uniform", 1, seed=SEED, chain=0, a=a, b=b, sweep=t, phase=phase)[0]


@pytest.fixture(scope="module", params=[64, 32])
def sweep(request):
    prec = request.param
    n_b = [25, 20, 15]
    G, K = 48, 3
    Y = bq.simulate(n_b, G, K, 99)
    o = po.Oracle(Y, n_b, K, seed=SEED, precision=prec, n_iter=20, n_burnin=10, n_unchanged=2)
    o.sweeps(4)
    t = o.iteration
    pre = o.state()
    pri = o.priors()
    o.sweeps(1)
    post = o.state()
    dbg = o.debug()
    return dict(prec=prec, Y=Y, n_b=n_b, G=G, K=K, t=t, pre=pre, post=post, pri=pri, d=dbg,
                batch=np.repeat(np.arange(len(n_b)), n_b))


FIX = 2.0 ** -33   # half the fixed-point resolution: each exact group sum carries at most this rounding


def tol(s, scale, groups=64):
    return (1e-9 if s["prec"] == 64 else 3e-6) * scale + groups * FIX


def mean(a, be, nu, de, W, batch):
    return np.exp(a[None, :] + be[W, :] + nu[batch, :] + de[:, None])


def test_subtype_logits_and_draw(sweep):
    s = sweep
    p, post = s["pre"], s["post"]
    X, b = p["X"], s["batch"]
    for k in range(1, s["K"]):
        mk = np.exp(p["alpha"][None] + p["beta"][k][None] + p["nu"][b] + p["delta"][:, None])
        m0 = np.exp(p["alpha"][None] + p["nu"][b] + p["delta"][:, None])
        terms = lnb(X, mk, p["phi"][b]) - lnb(X, m0, p["phi"][b])
        ref = terms.sum(1)
        assert np.all(np.abs(s["d"]["w_logit"][:, k] - ref) <= tol(s, 1 + np.abs(terms).sum(1)))
    lp = s["d"]["w_logit"] + np.log(p["pi"][b])
    w = np.exp(lp - lp.max(1, keepdims=True))
    for i in range(len(b)):
        tgt = u(i, 0, s["t"], PH["W"]) * w[i].sum()
        k = int(np.searchsorted(np.cumsum(w[i]), tgt))
        assert post["W"][i] == min(k, s["K"] - 1)


def test_alpha_sums_and_decisions(sweep):
    s = sweep
    p, post, d, b = s["pre"], s["post"], s["d"], s["batch"]
    X = post["X"]
    mu = mean(p["alpha"], p["beta"], p["nu"], p["delta"], post["W"], b)
    dl = d["alpha_prop"] - p["alpha"]
    terms = lnb(X, mu * np.exp(dl)[None], p["phi"][b]) - lnb(X, mu, p["phi"][b])
    assert np.all(np.abs(d["alpha_sum"] - terms.sum(0)) <= tol(s, 1 + np.abs(terms).sum(0)))
    for g in range(s["G"]):
        lr = d["alpha_sum"][g] + (st.norm.logpdf(d["alpha_prop"][g], s["pri"]["mu_alpha"][g], np.sqrt(5))
                                  - st.norm.logpdf(p["alpha"][g], s["pri"]["mu_alpha"][g], np.sqrt(5)))
        acc = np.log(u(g, 0, s["t"], PH["ALPHA_A"])) <= lr
        assert post["alpha"][g] == (d["alpha_prop"][g] if acc else p["alpha"][g])


def test_gamma_sums_and_decisions(sweep):
    s = sweep
    p, post, d, b = s["pre"], s["post"], s["d"], s["batch"]
    B = len(s["n_b"])
    X, Z = post["X"], post["Z"]
    sp = lambda v: np.logaddexp(0.0, v)
    for bb in range(B):
        x, z = X[b == bb].ravel(), Z[b == bb].ravel()
        g0, g1 = p["gamma"][0, bb], p["gamma"][1, bb]
        g0p, g1p = d["gamma_prop"][0, bb], d["gamma_prop"][1, bb]
        s0 = ((g0p - g0) * z - sp(g0p + g1 * x) + sp(g0 + g1 * x)).sum()
        assert abs(d["gamma_sum"][0, bb] - s0) < 1e-8 * (1 + np.abs(sp(g0 + g1 * x)).sum())
        lr0 = s0 + (st.norm.logpdf(g0p, 0, np.sqrt(3)) - st.norm.logpdf(g0, 0, np.sqrt(3)))
        acc0 = np.log(u(bb, 0, s["t"], PH["GAMMA0_A"])) <= lr0
        assert post["gamma"][0, bb] == (g0p if acc0 else g0)
        g0n = post["gamma"][0, bb]
        s1 = ((g1p - g1) * x * z - sp(g0n + g1p * x) + sp(g0n + g1 * x)).sum()
        assert abs(d["gamma_sum"][1, bb] - s1) < 1e-8 * (1 + np.abs(sp(g0n + g1 * x)).sum())


def test_L_decisions(sweep):
    s = sweep
    p, post = s["pre"], s["post"]
    pp, tau0 = p["ptau"]
    for k in range(1, s["K"]):
        for g in range(s["G"]):
            be = p["beta"][k, g]
            lo = (np.log(pp) + st.norm.logpdf(be, 0, 10.0) - np.log1p(-pp) - st.norm.logpdf(be, 0, np.sqrt(tau0)))
            assert post["L"][k, g] == int(u(g, k, s["t"], PH["L"]) > sc.expit(-lo))


def test_beta_nu_sums_and_decisions(sweep):
    s = sweep
    p, post, d, b = s["pre"], s["post"], s["d"], s["batch"]
    X, W, phi = post["X"], post["W"], p["phi"][b]
    tau0 = post["ptau"][1]
    for k in range(1, s["K"]):
        rows = W == k
        mu = mean(post["alpha"], p["beta"], p["nu"], p["delta"], W, b)[rows]
        dl = d["beta_prop"][k] - p["beta"][k]
        terms = lnb(X[rows], mu * np.exp(dl)[None], phi[rows]) - lnb(X[rows], mu, phi[rows])
        assert np.all(np.abs(d["beta_sum"][k] - terms.sum(0)) <= tol(s, 1 + np.abs(terms).sum(0)))
        for g in range(s["G"]):
            v = 100.0 if post["L"][k, g] else tau0
            lr = d["beta_sum"][k, g] + (st.norm.logpdf(d["beta_prop"][k, g], 0, np.sqrt(v))
                                        - st.norm.logpdf(p["beta"][k, g], 0, np.sqrt(v)))
            acc = np.log(u(g, k, s["t"], PH["BETA_A"])) <= lr
            assert post["beta"][k, g] == (d["beta_prop"][k, g] if acc else p["beta"][k, g])
    for bb in range(1, len(s["n_b"])):
        rows = b == bb
        mu = mean(post["alpha"], post["beta"], p["nu"], p["delta"], W, b)[rows]
        dl = d["nu_prop"][bb] - p["nu"][bb]
        terms = lnb(X[rows], mu * np.exp(dl)[None], phi[rows]) - lnb(X[rows], mu, phi[rows])
        # the nu term given an accepted beta is a difference of two log1p of size (phi+x)|dbeta + dnu|
        dlb = np.abs(d["beta_prop"] - p["beta"])[W[rows]]
        comp = ((phi[rows] + X[rows]) * (dlb + np.abs(dl)[None])).sum(0)
        assert np.all(np.abs(d["nu_sum"][bb] - terms.sum(0)) <= tol(s, 1 + comp + np.abs(terms).sum(0)))
        for g in range(s["G"]):
            m = s["pri"]["mu_nu"][bb]
            lr = d["nu_sum"][bb, g] + (st.norm.logpdf(d["nu_prop"][bb, g], m, np.sqrt(5))
                                       - st.norm.logpdf(p["nu"][bb, g], m, np.sqrt(5)))
            acc = np.log(u(g, bb, s["t"], PH["NU_A"])) <= lr
            assert post["nu"][bb, g] == (d["nu_prop"][bb, g] if acc else p["nu"][bb, g])


def test_delta_sums_and_decisions(sweep):
    s = sweep
    p, post, d, b = s["pre"], s["post"], s["d"], s["batch"]
    X, W = post["X"], post["W"]
    mu = mean(post["alpha"], post["beta"], post["nu"], p["delta"], W, b)
    dl = d["delta_prop"] - p["delta"]
    terms = lnb(X, mu * np.exp(dl)[:, None], p["phi"][b]) - lnb(X, mu, p["phi"][b])
    assert np.all(np.abs(d["delta_sum"] - terms.sum(1)) <= tol(s, 1 + np.abs(terms).sum(1)))
    first = np.concatenate([[0], np.cumsum(s["n_b"])[:-1]])
    assert np.all(post["delta"][first] == 0.0)
    for i in range(len(b)):
        if i in first:
            continue
        m = s["pri"]["mu_delta"][i]
        lr = d["delta_sum"][i] + (st.norm.logpdf(d["delta_prop"][i], m, np.sqrt(5))
                                  - st.norm.logpdf(p["delta"][i], m, np.sqrt(5)))
        acc = np.log(u(i, 0, s["t"], PH["DELTA_A"])) <= lr
        assert post["delta"][i] == (d["delta_prop"][i] if acc else p["delta"][i])


def test_phi_sums_and_decisions(sweep):
    s = sweep
    p, post, d, b = s["pre"], s["post"], s["d"], s["batch"]
    X, W = post["X"], post["W"]
    mu = mean(post["alpha"], post["beta"], post["nu"], post["delta"], W, b)
    for bb in range(len(s["n_b"])):
        rows = b == bb
        f, fp = p["phi"][bb], d["phi_prop"][bb]
        terms = lnb(X[rows], mu[rows], fp[None]) - lnb(X[rows], mu[rows], f[None])
        x = X[rows]
        # magnitude of the cancelling components of each term (x log(phi'/phi) against (phi+x) log1p(..))
        comp = (x + 1) * np.abs(np.log(fp / f))[None] + np.abs(fp - f)[None] * (np.log1p(x) + np.log1p(mu[rows]) + 1)
        scale = 1 + comp.sum(0) + np.abs(terms).sum(0)
        assert np.all(np.abs(d["phi_sum"][bb] - terms.sum(0)) <= tol(s, scale))
        prior = st.gamma(1.0, scale=10.0)
        for g in range(s["G"]):
            lr = d["phi_sum"][bb, g] + (prior.logpdf(fp[g]) - prior.logpdf(f[g])
                                        + st.gamma(fp[g]).logpdf(f[g]) - st.gamma(f[g]).logpdf(fp[g]))
            acc = fp[g] > 0 and np.log(u(g, bb, s["t"], PH["PHI_A"])) <= lr
            assert post["phi"][bb, g] == (fp[g] if acc else f[g])


def test_pi_is_a_distribution(sweep):
    pi = sweep["post"]["pi"]
    assert np.all(pi > 0) and np.allclose(pi.sum(1), 1.0, atol=1e-14)


def test_observed_counts_are_never_imputed(sweep):
    s = sweep
    Y = s["Y"]
    X, Z = s["post"]["X"], s["post"]["Z"]
    assert np.array_equal(X[Y > 0], Y[Y > 0]) and not Z[Y > 0].any()
    assert np.all(X[(Y == 0) & (Z == 0)] == 0)


def test_thread_count_does_not_change_the_chain():
    n_b = [30, 20]
    Y = bq.simulate(n_b, 70, 3, 5)
    runs = []
    for th in (1, 5):
        o = po.Oracle(Y, n_b, 3, seed=8, precision=32, n_iter=20, n_burnin=5, n_unchanged=2, threads=th)
        o.sweeps(8)
        runs.append(o.state())
    po.load().orc_set_threads(8)
    for k in ("S", "W", "L"):
        assert np.array_equal(runs[0][k], runs[1][k])
    for k in ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi", "ptau"):
        assert np.array_equal(runs[0][k], runs[1][k])


def test_tile_rows_change_only_rounding():
    n_b = [30, 20]
    Y = bq.simulate(n_b, 70, 3, 5)
    st_ = []
    for tr in (64, 7):
        o = po.Oracle(Y, n_b, 3, seed=8, precision=64, n_iter=20, n_burnin=5, n_unchanged=2, tile_rows=tr)
        o.sweeps(2)
        st_.append(o.state())
    assert np.allclose(st_[0]["alpha"], st_[1]["alpha"], rtol=1e-12)


@pytest.mark.slow
def test_recovers_synthetic_truth():
    # A single short chain can settle in a mode with two subtypes merged or one
    # emptied (the posterior is multimodal); the chain of lowest BIC among eight
    # recovers the truth, and every chain that missed it has a higher BIC.
    from sklearn.metrics import adjusted_rand_score
    n_b = [60, 50, 40]
    Y, W, X = bq.simulate(n_b, 300, 3, 2024, want_truth=True)
    runs = []
    for chain in range(8):
        o = po.Oracle(Y, n_b, 3, seed=13579, chain=chain, precision=32, n_iter=300, n_burnin=150,
                      n_unchanged=50)
        o.sweeps(300)
        post = o.posterior()
        runs.append((post["bic"], adjusted_rand_score(W, post["W"]), o))
    bic, ari, o = min(runs, key=lambda r: r[0])
    assert ari > 0.95
    assert all(r[0] > bic for r in runs if r[1] < 0.9)
    s = o.state()
    # imputed true counts at the dropout positions track the hidden truth
    zeros = Y == 0
    assert abs(s["X"][zeros].mean() - X[zeros].mean()) < 0.25 * X[zeros].mean()


@pytest.fixture(scope="module")
def wide_sweep():
    # more subtypes than one register group of the subtype draw (16) holds
    n_b = [45, 44]
    G, K = 40, 40
    Y = bq.simulate(n_b, G, K, 98)
    o = po.Oracle(Y, n_b, K, seed=SEED, precision=64, n_iter=20, n_burnin=10, n_unchanged=2)
    o.sweeps(3)
    t = o.iteration
    pre = o.state()
    o.sweeps(1)
    post = o.state()
    dbg = o.debug()
    return dict(prec=64, Y=Y, n_b=n_b, G=G, K=K, t=t, pre=pre, post=post, d=dbg,
                batch=np.repeat(np.arange(len(n_b)), n_b))


def test_subtype_logits_draw_and_pi_with_forty_subtypes(wide_sweep):
    test_subtype_logits_and_draw(wide_sweep)
    test_pi_is_a_distribution(wide_sweep)
    assert wide_sweep["post"]["W"].max() < 40 and len(np.unique(wide_sweep["post"]["W"])) > 16
