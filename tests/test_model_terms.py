"""The model terms of include/busseq/model.h pinned against scipy: every
per-element log-likelihood ratio is a difference of negative-binomial log-pmfs,
every per-parameter MH term a difference of prior and proposal log-densities
(BUSseq_gpu.cu:578-600, 625-632, 640-642, 696-707, 730-737, 763-778, 806-818),
the BIC element a marginal likelihood (cu:967-1007), and the X/Z update leaves
the exact conditional law of (X, Z) given Y = 0 invariant (cu:484-512)."""
import numpy as np
import pytest
import scipy.special as sc
import scipy.stats as st

from oracle import pyoracle as po

RNG = np.random.default_rng(77)


def lnb(x, mu, phi):
    return st.nbinom.logpmf(x, phi, phi / (phi + mu))


def sample_elements(n=3000):
    x = np.concatenate([np.zeros(n // 4), RNG.integers(1, 40, n // 2), RNG.integers(40, 3000, n - n // 4 - n // 2)])
    mu = np.exp(RNG.uniform(-3, 6, n))
    phi = np.exp(RNG.uniform(np.log(0.1), np.log(80), n))
    dlt = RNG.normal(0, 0.1, n)
    return x.astype(float), mu, phi, dlt


@pytest.mark.parametrize("prec", [64, 32])
def test_random_walk_term_is_nb_logpmf_difference(prec):
    x, mu, phi, dlt = sample_elements()
    if prec == 32:
        mu, phi, dlt = (a.astype(np.float32).astype(float) for a in (mu, phi, dlt))
    r = mu / (phi + mu)
    args = np.stack([x, phi, r, dlt, np.expm1(dlt)], 1)
    got = po.terms("tmh", prec, args)
    ref = lnb(x, mu * np.exp(dlt), phi) - lnb(x, mu, phi)
    scale = 1.0 + (phi + x) * np.abs(dlt) + x * np.abs(dlt)
    tol = 1e-9 if prec == 64 else 2e-6
    assert (np.abs(got - ref) / scale).max() < tol
    # planted mistake: a dropped (phi + x) factor must be detected
    bad = x * dlt - np.log1p(r * np.expm1(dlt))
    assert (np.abs(bad - ref) / scale).max() > 100 * tol


@pytest.mark.parametrize("prec", [64, 32])
def test_subtype_logit_term(prec):
    x, mu0, phi, _ = sample_elements()
    beta = RNG.normal(0, 2, len(x))
    if prec == 32:
        mu0, phi, beta = (a.astype(np.float32).astype(float) for a in (mu0, phi, beta))
    r0 = mu0 / (phi + mu0)
    got = po.terms("tw", prec, np.stack([x, phi, r0, beta, np.expm1(beta)], 1))
    ref = lnb(x, mu0 * np.exp(beta), phi) - lnb(x, mu0, phi)
    scale = 1.0 + (phi + x) * np.abs(beta) + x * np.abs(beta)
    assert (np.abs(got - ref) / scale).max() < (1e-9 if prec == 64 else 3e-6)


@pytest.mark.parametrize("prec", [64, 32])
def test_phi_term_is_nb_logpmf_difference(prec):
    x, mu, phi, _ = sample_elements()
    phin = phi * np.exp(RNG.normal(0, 0.3, len(x)))
    if prec == 32:
        mu = mu.astype(np.float32).astype(float)
    got = po.terms("tphi", prec, np.stack([x, phi, phin, mu], 1))
    ref = lnb(x, mu, phin) - lnb(x, mu, phi)
    # magnitude of the components that cancel: x log(phi'/phi) against (phi+x) log1p(...)
    scale = 1.0 + (x + 1) * np.abs(np.log(phin / phi)) + np.abs(phin - phi) * (np.log1p(x) + np.log1p(mu) + 1)
    assert (np.abs(got - ref) / scale).max() < (1e-9 if prec == 64 else 2e-6)


def test_gamma1_prior_and_proposal():
    g = np.exp(RNG.uniform(-4, 1, 500))
    gn = g * np.exp(RNG.normal(0, 0.3, 500))
    got = po.terms("lr_gamma1", 64, np.stack([g, gn], 1))
    prior = st.gamma(0.001, scale=1 / 0.01)
    ref = (prior.logpdf(gn) - prior.logpdf(g)
           + st.gamma(10 * gn, scale=0.1).logpdf(g) - st.gamma(10 * g, scale=0.1).logpdf(gn))
    assert np.abs(got - ref).max() < 1e-9 * (1 + np.abs(ref).max())


def test_phi_prior_and_proposal():
    f = np.exp(RNG.uniform(-2, 4, 500))
    fn = st.gamma(f).rvs(random_state=1)
    got = po.terms("lr_phi", 64, np.stack([f, fn], 1))
    prior = st.gamma(1.0, scale=1 / 0.1)
    ref = prior.logpdf(fn) - prior.logpdf(f) + st.gamma(fn).logpdf(f) - st.gamma(f).logpdf(fn)
    assert np.abs(got - ref).max() < 1e-9 * (1 + np.abs(ref).max())


def test_spike_slab_log_odds():
    p = RNG.uniform(0.01, 0.99, 500)
    tau0 = np.exp(RNG.uniform(-7, -2, 500))
    beta = RNG.normal(0, 0.3, 500)
    got = po.terms("logodds_L", 64, np.stack([p, tau0, beta], 1))
    ref = (np.log(p) + st.norm.logpdf(beta, 0, np.sqrt(100.0))
           - np.log1p(-p) - st.norm.logpdf(beta, 0, np.sqrt(tau0)))
    assert np.abs(got - ref).max() < 1e-9 * (1 + np.abs(ref).max())


def test_gaussian_priors():
    old, new, m = RNG.normal(0, 2, (3, 500))
    s2 = np.exp(RNG.uniform(-3, 3, 500))
    got = po.terms("lr_normal", 64, np.stack([old, new, m, s2], 1))
    ref = st.norm.logpdf(new, m, np.sqrt(s2)) - st.norm.logpdf(old, m, np.sqrt(s2))
    assert np.abs(got - ref).max() < 1e-10 * (1 + np.abs(ref).max())
    L = RNG.integers(0, 2, 500).astype(float)
    got = po.terms("lr_beta", 64, np.stack([old, new, L, s2], 1))
    v = np.where(L > 0, 100.0, s2)
    ref = st.norm.logpdf(new, 0, np.sqrt(v)) - st.norm.logpdf(old, 0, np.sqrt(v))
    assert np.abs(got - ref).max() < 1e-10 * (1 + np.abs(ref).max())


def bic_elem_ref(a, bk, nb, dl, g0, g1, phi, y):
    """numpy restatement of cu:967-1007 (dropout marginalised, floor(3 mu) truncation, -250 clip)"""
    mu = np.exp(a + bk + nb + dl)
    p = mu / (mu + phi)
    logp = -250.0 if p < np.exp(-250.0) else np.log(p)
    log1mp = -250.0 if 1 - p < np.exp(-250.0) else np.log(1 - p)
    lnbx = lambda x: sc.gammaln(phi + x) - sc.gammaln(1 + x) - sc.gammaln(phi) + x * logp + phi * log1mp
    if y > 0:
        return -np.logaddexp(0, g0 + g1 * y) + lnbx(y)
    xm = int(3 * mu)
    terms = [phi * log1mp] + [g0 + g1 * x - np.logaddexp(0, g0 + g1 * x) + lnbx(x) for x in range(1, xm)]
    return sc.logsumexp(terms)


def test_bic_element_vs_restatement():
    rows = []
    for _ in range(300):
        a, bk, nb, dl = RNG.normal(1.5, 1.0), RNG.normal(0, 1), RNG.normal(0, 0.5), RNG.normal(0, 0.2)
        g0, g1 = RNG.uniform(-2, 0), RNG.uniform(-0.3, -0.1)
        phi = np.exp(RNG.uniform(-1, 3))
        y = 0 if RNG.random() < 0.4 else int(RNG.integers(1, 200))
        rows.append([a, bk, nb, dl, g0, g1, phi, y])
    rows = np.array(rows)
    got = po.terms("bic_elem", 64, rows)
    ref = np.array([bic_elem_ref(*r[:7], int(r[7])) for r in rows])
    assert np.abs(got - ref).max() < 1e-8 * (1 + np.abs(ref).max())


@pytest.mark.parametrize("mu,phi,g0,g1", [(3.0, 5.0, -1.0, -0.2), (0.4, 0.8, 0.3, -0.1), (40.0, 2.0, -2.0, -0.05)])
def test_xz_update_preserves_the_conditional_law(mu, phi, g0, g1):
    n, nx = 400000, 400
    counts = po.xz_chain(mu, phi, g0, g1, seed=12, n=n, nx=nx)
    x = np.arange(nx)
    nb = st.nbinom.pmf(x, phi, phi / (phi + mu))
    target = np.zeros((nx, 2))
    target[0, 0] = nb[0] * (1 - sc.expit(g0))
    target[0, 1] = nb[0] * sc.expit(g0)
    target[1:, 1] = nb[1:] * sc.expit(g0 + g1 * x[1:])
    target /= target.sum()
    emp = counts / counts.sum()
    tv = 0.5 * np.abs(emp - target).sum()
    assert tv < 0.01, tv
    assert counts[1:, 0].sum() == 0            # X > 0 with Z = 0 is impossible when Y = 0
    # planted mistake: forgetting the dropout factor (sampling X from the NB alone) is detected
    wrong = np.zeros_like(target)
    wrong[:, 1] = nb
    wrong /= wrong.sum()
    assert 0.5 * np.abs(emp - wrong).sum() > 0.05


@pytest.mark.timeout(60)
def test_bic_element_of_a_nan_mean_returns_at_once():
    # a chain whose parameters left the finite range: the zero-count element sums no
    # terms (instead of running its loop to the 2e9 cap) and comes out NaN
    rows = np.array([[np.nan, 0.0, 0.0, 0.0, -1.0, -0.2, 2.0, 0.0],
                     [np.nan, 0.0, 0.0, 0.0, -1.0, -0.2, 2.0, 3.0],
                     [1.0, 0.0, 0.0, 0.0, -1.0, -0.2, np.nan, 0.0]])
    got = po.terms("bic_elem", 64, rows)
    assert np.isnan(got).all()
