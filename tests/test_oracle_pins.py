"""Independent pins of the oracle's initialisation, pi and p/tau0 draws,
posterior summaries and BIC, each with a planted mistake that it must catch.

Every check below is a restatement of the reference (BUSseq_gpu.cu, cited
per check) written with numpy/scipy from the reference's own formulas and
constants, not from the oracle's code.  Each is run twice: on the oracle as
built (it must pass) and on the oracle with one plausible mistake planted
(orc_set_mutation; it must fail).  The GPU sampler equals the oracle bit for
bit (tests/test_gpu_*.py), so these checks pin it too.

  init       cu:369-482, 1163-1244: library sizes, pi0, W0 with its Philox
             uniform, delta0, raw means (0 for an empty group), mu_nu, nu0,
             beta0, alpha0, L0 with its Philox uniform, phi0, gamma0, p0, tau0
  pi         cu:856-890: the marginals of pi_b ~ Dirichlet(n_b. + 2) are
             Beta(n_bk + 2, sum_j (n_bj + 2) - n_bk - 2), KS over 4000 draws
  p, tau0    cu:650-668: p ~ Beta(sum L + 1, G(K-1) - sum L + 3) and
             tau0 ~ InvGamma(2 + n0/2, 0.01 + sum_{L=0} beta^2 / 2), KS
  posterior  cu:938-959, 1557-1608: means of the recorded traces, P(L = 1),
             the W mode with ties to the lowest subtype; the traces equal the
             state after each recorded sweep
  BIC        cu:967-1024, 1815-1833: the per-count likelihood with dropout
             marginalised (sum over x < floor(3 mu) for y = 0, log p clipped
             at -250), the pi-weighted mixture over subtypes and the penalty
             log(GN) ((B+G)K + 2B + G(2B-1) + N - B)
"""
import numpy as np
import pytest
import scipy.special as sc
import scipy.stats as st

import busseq_gpu_amd as bq
from oracle import pyoracle as po

SEED = 2718
SWEEP_INIT = 0xFFFFFFFF
PH_INIT_W, PH_INIT_L, PH_INIT_P = 1, 2, 3
# hyperparameters of the reference, cu:12-32
P_INIT_BOUND, TAU0_INIT, GAMMA0_INIT, GAMMA1_INIT, PHI_INIT = 0.5, 0.005, 0.0, -0.3, 5.0
XI_PI, A_P, B_P, A_TAU0, B_TAU0, TAU1 = 2.0, 1.0, 3.0, 2.0, 0.01, 100.0


@pytest.fixture(autouse=True)
def no_mutation():
    po.set_mutation(0)
    yield
    po.set_mutation(0)


def uni(a, b, sweep, phase, chain=0):
    return po.stream("uniform", 1, seed=SEED, chain=chain, a=a, b=b, sweep=sweep, phase=phase)[0]


# ---------------------------------------------------------------------------
# initialisation
# ---------------------------------------------------------------------------
def init_data():
    n_b = [7, 5, 6]
    G, K = 9, 3
    Y = bq.simulate(n_b, G, K, 314)
    Y[:, 4] = 0                      # a gene never observed
    return Y, n_b, K


def init_reference(Y, n_b, K):
    """cu:1163-1244 restated: the initial state and the prior means"""
    N, G = Y.shape
    B = len(n_b)
    first = np.concatenate([[0], np.cumsum(n_b)[:-1]])
    batch = np.repeat(np.arange(B), n_b)
    lib = Y.sum(1).astype(np.float64)                                    # cu:1204-1205
    pi0 = np.tile((np.arange(K) + 1) / (K * (K + 1) / 2.0), (B, 1))      # cu:392-397
    W = np.zeros(N, np.int64)
    for i in range(N):                                                   # cu:425-432
        u = uni(i, 0, SWEEP_INIT, PH_INIT_W)
        k, tmp = 0, pi0[batch[i], 0]
        while u > tmp and k < K - 1:
            k += 1
            tmp += pi0[batch[i], k]
        W[i] = k
    delta = np.log(lib) - np.log(lib[first[batch]])                       # cu:435-440
    lg = np.log1p(Y / np.exp(delta)[:, None])                            # cu:369-382
    raw = np.zeros((B, K, G))
    empty = []
    for b in range(B):
        for k in range(K):
            rows = (batch == b) & (W == k)
            if rows.any():
                raw[b, k] = lg[rows].mean(0)
            else:
                empty.append((b, k))                                     # cu:289-292: mean 0
    alpha = raw[0, 0].copy()                                             # cu:1227
    beta = raw[0] - raw[0, 0][None]                                      # cu:456
    beta[0] = 0.0
    nu = (raw.sum(1) - raw[0].sum(0)[None]) / K                          # cu:399-417
    nu[0] = 0.0
    bm = np.array([lg[batch == b].mean() for b in range(B)])            # cu:470-482, 1232-1241
    mu_nu = bm - bm[0]
    p0 = uni(0, 0, SWEEP_INIT, PH_INIT_P) * P_INIT_BOUND                 # cu:1172-1173
    lr_base = np.log(p0 / (1 - p0)) + (np.log(TAU0_INIT) - np.log(TAU1)) / 2
    L = np.zeros((K, G), np.int64)
    for k in range(1, K):                                                # cu:445-462
        for g in range(G):
            lr = lr_base + beta[k, g] ** 2 * (1 / TAU0_INIT - 1 / TAU1) / 2
            L[k, g] = uni(g, k, SWEEP_INIT, PH_INIT_L) > 1.0 / (1.0 + np.exp(lr))
    return dict(W=W, delta=delta, alpha=alpha, beta=beta, nu=nu, L=L, pi=pi0,
                phi=np.full((B, G), PHI_INIT), gamma=np.array([[GAMMA0_INIT] * B, [GAMMA1_INIT] * B]),
                ptau=np.array([p0, TAU0_INIT]), X=Y.copy(), Z=np.zeros_like(Y),
                mu_alpha=alpha.copy(), mu_nu=mu_nu, mu_delta=delta.copy(), empty=empty)


def init_mismatches(Y, n_b, K):
    o = po.Oracle(Y, n_b, K, seed=SEED, n_iter=10, n_burnin=5, n_unchanged=2)
    s, pri = o.state(), o.priors()
    ref = init_reference(Y, n_b, K)
    bad = []
    for key in ("W", "L", "X", "Z"):
        if not np.array_equal(s[key], ref[key]):
            bad.append(key)
    # fixed-point sums round each term to 2^-33: means are exact to ~1e-9
    for key, tol in (("delta", 1e-13), ("alpha", 1e-9), ("beta", 1e-9), ("nu", 1e-9), ("pi", 1e-15),
                     ("phi", 0), ("gamma", 0), ("ptau", 1e-15)):
        if not np.allclose(s[key], ref[key], rtol=0, atol=tol):
            bad.append(key)
    for key, tol in (("mu_alpha", 1e-9), ("mu_nu", 1e-9), ("mu_delta", 1e-13)):
        if not np.allclose(pri[key], ref[key], rtol=0, atol=tol):
            bad.append(key)
    return bad, ref


def test_init_matches_the_reference_restatement():
    Y, n_b, K = init_data()
    bad, ref = init_mismatches(Y, n_b, K)
    assert bad == []
    # the data exercise what the rules distinguish
    assert ref["empty"], "an empty (batch, subtype) group"
    assert np.all(np.abs(ref["mu_nu"][1:]) > 1e-3) and ref["L"][1:].any() and not ref["L"][1:].all()


@pytest.mark.parametrize("mutation, key", [(3, "mu_nu"), (8, "alpha"), (9, "delta")])
def test_init_check_catches_a_planted_mistake(mutation, key):
    Y, n_b, K = init_data()
    po.set_mutation(mutation)
    bad, _ = init_mismatches(Y, n_b, K)
    assert key in bad, bad


# ---------------------------------------------------------------------------
# pi_b ~ Dirichlet(n_b. + xi)
# ---------------------------------------------------------------------------
def pi_min_pvalue(draws=4000):
    n_b, K = [12, 9], 3
    Y = bq.simulate(n_b, 6, K, 17)
    W = np.array([0] * 10 + [1] * 2 + [2] * 8 + [1] * 1, np.int32)   # counts (10, 2, 0) and (0, 1, 8)
    cnt = np.array([[10, 2, 0], [0, 1, 8]])
    o = po.Oracle(Y, n_b, K, seed=SEED, n_iter=10, n_burnin=5, n_unchanged=2)
    o.set_state({"W": W})
    pis = np.zeros((draws, 2, K))
    for t in range(draws):
        o.step("pi", t)
        pis[t] = o.state()["pi"]
    assert np.allclose(pis.sum(2), 1.0, atol=1e-14)
    a = cnt + XI_PI
    pv = [st.kstest(pis[:, b, k], st.beta(a[b, k], a[b].sum() - a[b, k]).cdf).pvalue
          for b in range(2) for k in range(K)]
    return min(pv)


def test_pi_draws_follow_the_dirichlet_conditional():
    assert pi_min_pvalue() > 1e-3


def test_pi_check_catches_a_planted_prior():
    po.set_mutation(1)
    assert pi_min_pvalue() < 1e-6


# ---------------------------------------------------------------------------
# p ~ Beta, tau0 ~ InvGamma
# ---------------------------------------------------------------------------
def ptau_pvalues(draws=4000):
    n_b, G, K = [4, 4], 4, 3
    Y = bq.simulate(n_b, G, K, 23)
    L = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0]], np.int32)
    beta = np.array([[0, 0, 0, 0], [1.5, 0.05, -0.02, 0.03], [0.01, -0.04, 2.0, 0.06]])
    o = po.Oracle(Y, n_b, K, seed=SEED, n_iter=10, n_burnin=5, n_unchanged=2)
    o.set_state({"L": L, "beta": beta})
    out = np.zeros((draws, 2))
    for t in range(draws):
        o.step("ptau", t)
        out[t] = o.state()["ptau"]
    n1 = L[1:].sum()
    n0 = G * (K - 1) - n1
    sb = (beta[1:][L[1:] == 0] ** 2).sum()
    p_law = st.beta(n1 + A_P, G * (K - 1) - n1 + B_P)
    tau_law = st.invgamma(A_TAU0 + n0 / 2.0, scale=B_TAU0 + sb / 2.0)
    return st.kstest(out[:, 0], p_law.cdf).pvalue, st.kstest(out[:, 1], tau_law.cdf).pvalue


def test_p_and_tau0_draws_follow_their_conditionals():
    pp, pt = ptau_pvalues()
    assert pp > 1e-3 and pt > 1e-3


@pytest.mark.parametrize("mutation, which", [(2, 0), (4, 1)])
def test_p_tau0_check_catches_a_planted_mistake(mutation, which):
    po.set_mutation(mutation)
    assert ptau_pvalues()[which] < 1e-6


# ---------------------------------------------------------------------------
# posterior summaries from the traces
# ---------------------------------------------------------------------------
def posterior_mismatches():
    n_b, G, K = [14, 12, 10], 6, 3          # few genes: subtypes change often, so W modes tie
    Y = bq.simulate(n_b, G, K, 41)
    n_iter, n_burnin = 12, 8
    o = po.Oracle(Y, n_b, K, seed=SEED, n_iter=n_iter, n_burnin=n_burnin, n_unchanged=2, keep_traces=True)
    o.sweeps(n_burnin)
    states = []
    for _ in range(n_iter - n_burnin):
        o.sweeps(1)
        states.append(o.state())
    tr, post = o.traces(), o.posterior()
    bad = []
    # every trace column is the state after its recorded sweep (cu:892-936)
    for r, s in enumerate(states):
        cols = {"gamma": s["gamma"].ravel(), "alpha": s["alpha"], "L": s["L"][1:].ravel(),
                "beta": s["beta"][1:].ravel(), "ptau": s["ptau"], "nu": s["nu"][1:].ravel(),
                "delta": s["delta"], "phi": s["phi"].ravel(), "W": s["W"], "pi": s["pi"].ravel()}
        for key, v in cols.items():
            if not np.array_equal(tr[key][:, r], v):
                bad.append(f"trace {key}")
    # means of the recorded iterations (cu:1557-1608); beta and nu carry a zero first row
    means = {"gamma": tr["gamma"].mean(1).reshape(2, -1), "alpha": tr["alpha"].mean(1),
             "beta": np.vstack([np.zeros(G), tr["beta"].mean(1).reshape(K - 1, G)]),
             "L": np.vstack([np.zeros(G), tr["L"].mean(1).reshape(K - 1, G)]),
             "ptau": tr["ptau"].mean(1), "nu": np.vstack([np.zeros(G), tr["nu"].mean(1).reshape(-1, G)]),
             "delta": tr["delta"].mean(1), "phi": tr["phi"].mean(1).reshape(-1, G),
             "pi": tr["pi"].mean(1).reshape(-1, K)}
    for key, v in means.items():
        if not np.allclose(post[key], v, rtol=1e-13, atol=1e-15):
            bad.append(key)
    # the W mode, ties to the lowest subtype (strict > in mode_on_gpu, cu:944-959)
    counts = np.stack([(tr["W"] == k).sum(1) for k in range(K)], 1)
    mode = counts.argmax(1)                                  # argmax returns the first maximum
    ties = (counts == counts.max(1, keepdims=True)).sum(1) > 1
    if not np.array_equal(post["W"], mode):
        bad.append("W")
    return bad, int(ties.sum())


def test_posterior_summaries_match_the_traces():
    bad, ties = posterior_mismatches()
    assert bad == []
    assert ties >= 3, "the data exercise the tie-break"


@pytest.mark.parametrize("mutation, key", [(5, "W"), (10, "delta")])
def test_posterior_check_catches_a_planted_mistake(mutation, key):
    po.set_mutation(mutation)
    bad, _ = posterior_mismatches()
    assert key in bad, bad


# ---------------------------------------------------------------------------
# BIC
# ---------------------------------------------------------------------------
def bic_reference(Y, n_b, K, p):
    """cu:967-1024 and 1815-1833 restated with scipy, from the posterior means p"""
    N, G = Y.shape
    B = len(n_b)
    batch = np.repeat(np.arange(B), n_b)
    g0, g1 = p["gamma"][0][batch][:, None], p["gamma"][1][batch][:, None]
    phi = p["phi"][batch]
    ll = np.zeros((N, K))
    lim = np.exp(-250.0)
    for k in range(K):
        mu = np.exp(p["alpha"][None] + p["beta"][k][None] + p["nu"][batch] + p["delta"][:, None])
        pb = mu / (mu + phi)
        logp = np.where(pb < lim, -250.0, np.log(pb))
        log1mp = np.where((pb >= lim) & (1 - pb < lim), -250.0, np.log1p(-pb))
        pos = -np.logaddexp(0.0, g0 + g1 * Y) + sc.gammaln(phi + Y) - sc.gammaln(1.0 + Y) - sc.gammaln(phi) \
            + Y * logp + phi * log1mp
        zero = phi * log1mp
        xmax = np.floor(3 * mu).astype(np.int64)
        for x in range(1, int(xmax.max())):
            lin = g0 + g1 * x
            tmp = lin - np.logaddexp(0.0, lin) + sc.gammaln(phi + x) - sc.gammaln(1.0 + x) - sc.gammaln(phi) \
                + x * logp + phi * log1mp
            zero = np.where(x < xmax, np.logaddexp(zero, tmp), zero)
        ll[:, k] = np.where(Y > 0, pos, zero).sum(1)
    mx = ll.max(1)
    lli = mx + np.log((p["pi"][batch] * np.exp(ll - mx[:, None])).sum(1))
    npar = (B + G) * K + 2 * B + G * (2 * B - 1) + N - B
    return -2.0 * lli.sum() + np.log(G * N) * npar


def bic_pair():
    n_b, G, K = [20, 15, 15], 30, 3
    Y = bq.simulate(n_b, G, K, 59)
    o = po.Oracle(Y, n_b, K, seed=SEED, n_iter=30, n_burnin=15, n_unchanged=5)
    o.sweeps(30)
    p = o.posterior()
    return p["bic"], bic_reference(Y, n_b, K, p)


def test_bic_matches_the_reference_restatement():
    got, ref = bic_pair()
    assert abs(got - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("mutation", [6, 7])
def test_bic_check_catches_a_planted_mistake(mutation):
    po.set_mutation(mutation)
    got, ref = bic_pair()
    assert abs(got - ref) > 1e-6 * abs(ref)
