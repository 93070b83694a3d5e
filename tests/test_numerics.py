"""Numerics contract (include/busseq/numerics.h) pinned against published
known-answer tests, 40-digit mpmath values (tests/golden/math_vectors.npz,
written by tests/golden/make_math_vectors.py) and scipy distributions."""
import os

import numpy as np
import pytest
import scipy.stats as st

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "math_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


# Random123 known-answer vectors for philox4x32-10 (kat_vectors of the Random123 distribution)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert po.philox(ctr, key).tolist() == expect


def test_stream_words_follow_the_counter():
    # the stream's words are the Philox blocks of counters (a, b, sweep, phase<<20 | j)
    w = po.stream("u32", 8, seed=11, chain=3, a=5, b=6, sweep=7, phase=9)
    b0 = po.philox([5, 6, 7, 9 << 20], [11, 3])
    b1 = po.philox([5, 6, 7, (9 << 20) + 1], [11, 3])
    assert w.astype(np.uint32).tolist() == b0.tolist() + b1.tolist()


def test_uniform_open_interval_and_resolution():
    u = po.stream("uniform", 200000, seed=1)
    assert u.min() > 0.0 and u.max() < 1.0
    assert st.kstest(u, "uniform").pvalue > 1e-3
    frac = u * 2.0 ** 32
    assert np.all(frac - np.floor(frac) == 0.5)   # centred in a 2^-32 cell
    w = po.stream("u32", 64, seed=1)
    assert np.array_equal(u[:64], (w + 0.5) * 2.0 ** -32)   # one Philox word per uniform


def ulps(x, ref, dtype=np.float64):
    ref = np.asarray(ref, dtype)
    sp = np.spacing(np.maximum(np.abs(ref), np.finfo(dtype).tiny)).astype(np.float64)
    return np.abs(np.asarray(x, np.float64) - ref.astype(np.float64)) / sp


def test_log_vs_mpmath(gold):
    assert ulps(po.math("log", gold["log_x"]), gold["log_y"]).max() <= 1.0


def test_log1p_vs_mpmath(gold):
    assert ulps(po.math("log1p", gold["log1p_x"]), gold["log1p_y"]).max() <= 2.0


def test_exp_vs_mpmath(gold):
    assert ulps(po.math("exp", gold["exp_x"]), gold["exp_y"]).max() <= 1.0


def test_expm1_vs_mpmath(gold):
    assert ulps(po.math("expm1", gold["expm1_x"]), gold["expm1_y"]).max() <= 4.0


def test_lgamma_vs_mpmath(gold):
    y = gold["lgamma_y"]
    err = np.abs(po.math("lgamma", gold["lgamma_x"]) - y) / np.maximum(np.abs(y), 1.0)
    # absolute error near the zeros at 1 and 2 comes from lgamma(z+n) - log(prod), ~3 ulp of lgamma(10)
    assert err.max() < 1e-14


def test_lgamma_diff_vs_mpmath(gold):
    z, d, y = gold["lgdiff_z"], gold["lgdiff_d"], gold["lgdiff_y"]
    e64 = np.abs(po.math("lgamma_diff64", z, d) - y) / np.maximum(np.abs(y), 1.0)
    assert e64.max() < 1e-13
    # single precision: arguments are first rounded to float; compare to the
    # exact difference at the rounded arguments within a few float ulps of the scale
    zf, df = z.astype(np.float32), d.astype(np.float32)
    import mpmath as mp
    ref = np.array([float(mp.loggamma(mp.mpf(float(a)) + mp.mpf(float(b))) - mp.loggamma(mp.mpf(float(a))))
                    for a, b in zip(zf[:400], df[:400])])
    e32 = np.abs(po.math("lgamma_diff32", zf[:400].astype(np.float64), df[:400].astype(np.float64)) - ref)
    assert (e32 / np.maximum(np.abs(ref), 1.0)).max() < 2e-6


def test_lgamma_diff32_on_the_phi_term_domain():
    # the phi term evaluates lgamma(phi' + x) - lgamma(phi + x) with x >= 1 and
    # phi' ~ Gamma(phi, 1): arguments above 1, where the fitted Stirling remainder
    # is used without a shift; checked against 40-digit mpmath at the float arguments
    import mpmath as mp
    rng = np.random.default_rng(11)
    n = 1500
    phi = np.exp(rng.uniform(np.log(0.02), np.log(60), n)).astype(np.float32)
    x = np.where(rng.random(n) < 0.5, rng.integers(1, 2001, n), rng.integers(1, 12, n)).astype(np.float32)
    phip = rng.gamma(phi.astype(np.float64), 1.0).astype(np.float32)
    z = (phi + x).astype(np.float32)
    d = (phip - phi).astype(np.float32)
    assert z.min() > 1.0
    got = po.math("lgamma_diff32", z.astype(np.float64), d.astype(np.float64))
    with mp.workdps(40):
        ref = np.array([float(mp.loggamma(mp.mpf(float(a)) + mp.mpf(float(b))) - mp.loggamma(mp.mpf(float(a))))
                        for a, b in zip(z, d)])
    assert (np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)).max() < 5e-7


def test_lgamma_diff64_on_the_phi_term_domain():
    # the double-precision phi term evaluates lgamma(phi' + x) - lgamma(phi + x) with
    # x >= 1: arguments above 1, where the fitted Stirling remainder (1 <= w < 10) and
    # the series (w >= 10) are used without a shift; 40-digit mpmath
    import mpmath as mp
    rng = np.random.default_rng(12)
    n = 2000
    phi = np.exp(rng.uniform(np.log(0.02), np.log(60), n))
    x = np.where(rng.random(n) < 0.5, rng.integers(1, 2001, n), rng.integers(1, 12, n)).astype(np.float64)
    z = phi + x
    d = rng.gamma(phi, 1.0) - phi
    got = po.math("lgamma_diff64", z, d)
    with mp.workdps(40):
        ref = np.array([float(mp.loggamma(mp.mpf(a) + mp.mpf(b)) - mp.loggamma(mp.mpf(a))) for a, b in zip(z, d)])
    assert (np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)).max() < 2e-15


def test_lgamma_diff64_across_the_fit_and_series_ranges():
    # the Stirling remainder switches from the fit to the series at w = 10: arguments
    # on both sides of it and at it, against 40-digit mpmath
    import mpmath as mp
    z = np.concatenate([np.linspace(1.0, 12.0, 1101), [np.nextafter(10.0, 0), 10.0, np.nextafter(10.0, 20)]])
    d = np.concatenate([np.full(1101, 0.75), [0.5, -0.5, 1e-9]])
    with mp.workdps(40):
        ref = np.array([float(mp.loggamma(mp.mpf(float(a)) + mp.mpf(float(b))) - mp.loggamma(mp.mpf(float(a))))
                        for a, b in zip(z, d)])
    err = np.abs(po.math("lgamma_diff64", z, d) - ref) / np.maximum(np.abs(ref), 1.0)
    assert err.max() < 2e-15


def test_log1p_mid_vs_mpmath():
    # the branch-free log1p of the double-precision element terms, on (-1, 1e12)
    import mpmath as mp
    rng = np.random.default_rng(5)
    y = np.concatenate([rng.uniform(-0.999999, 3, 1500), rng.uniform(-1e-8, 1e-8, 200), -1 + np.exp(rng.uniform(-30, -1, 200)),
                        np.exp(rng.uniform(-40, 27, 400)), [0.0, 0.41421356237309515, -0.29289321881345248]])
    with mp.workdps(40):
        ref = np.array([float(mp.log1p(mp.mpf(float(v)))) for v in y])
    assert ulps(po.math("log1p_mid", y), ref).max() <= 2.0


def test_single_precision_logs_vs_mpmath(gold):
    x = gold["log_x"].astype(np.float32)
    x = x[np.isfinite(x) & (x > 0)]
    import mpmath as mp
    ref = np.array([float(mp.log(mp.mpf(float(v)))) for v in x])
    assert ulps(po.math("logf", x.astype(np.float64)), ref, np.float32).max() <= 1.0
    y = gold["log1p_x"].astype(np.float32)
    y = y[y > -1]
    ref = np.array([float(mp.log1p(mp.mpf(float(v)))) for v in y])
    assert ulps(po.math("log1pf", y.astype(np.float64)), ref, np.float32).max() <= 3.0


def test_single_precision_log1p_over_its_reduction_ranges():
    # every exponent k of 1 + y from -24 to 127, |y| tiny (f = y exactly), and both
    # sides of the switch to the rare path at 1 + y = sqrt(2) 2^125
    import mpmath as mp
    rng = np.random.default_rng(9)
    y = np.concatenate([
        -1 + np.exp2(-rng.uniform(0, 24, 400)),
        np.exp2(rng.uniform(-40, 127.9, 800)),
        -np.exp2(rng.uniform(-40, -0.1, 300)),
        np.float32(2 ** 125.5) * (1 + rng.uniform(-1e-6, 1e-6, 50)),
    ]).astype(np.float32)
    y = y[np.isfinite(y) & (y > -1)]
    ref = np.array([float(mp.log1p(mp.mpf(float(v)))) for v in y])
    assert ulps(po.math("log1pf", y.astype(np.float64)), ref, np.float32).max() <= 2.0


def test_single_precision_log1p_mid_over_its_reduction_ranges():
    # the branch-free log1pf of the subtype and random-walk terms, degree-7
    # polynomial: within 2 ulp, as log1pf, for every exponent k of 1 + y from -24
    # to 125, and y tiny (f = y exactly); mean error at most 0.35 ulp
    import mpmath as mp
    rng = np.random.default_rng(11)
    y = np.concatenate([
        -1 + np.exp2(-rng.uniform(0, 24, 400)),
        np.exp2(rng.uniform(-40, 125.4, 1200)),
        -np.exp2(rng.uniform(-40, -0.1, 600)),
        rng.uniform(np.sqrt(0.5) - 1, np.sqrt(2) - 1, 3000),
    ]).astype(np.float32)
    y = y[np.isfinite(y) & (y > -1)]
    ref = np.array([float(mp.log1p(mp.mpf(float(v)))) for v in y])
    err = ulps(po.math("log1pf_mid", y.astype(np.float64)), ref, np.float32)
    assert err.max() <= 2.0 and err.mean() <= 0.35
    # planted mistake: the degree-7 polynomial with its last coefficient dropped
    # (a degree-6 truncation) is far outside the bound on the same inputs
    f = np.float32(y[-3000:])
    q6 = np.polyval(np.float32([-1.41752481460571289063e-1, 1.48381710052490234375e-1,
                                -1.65869489312171936035e-1, 1.99716746807098388672e-1,
                                -2.50011831521987915039e-1, 3.33336770534515380859e-1,
                                -4.99999970197677612305e-1]), f.astype(np.float64))
    assert ulps(f + f * f * q6, ref[-3000:], np.float32).max() > 4.0


def test_ndtri_vs_mpmath(gold):
    y = gold["ndtri_y"]
    err = np.abs(po.math("ndtri", gold["ndtri_x"]) - y) / np.maximum(np.abs(y), 1.0)
    assert err.max() < 1e-14


def test_special_values():
    inf = np.inf
    r = po.math("log", [0.0, -1.0, inf, np.nan])
    assert r[0] == -inf and np.isnan(r[1]) and r[2] == inf and np.isnan(r[3])
    r = po.math("log1p", [-1.0, -2.0, inf])
    assert r[0] == -inf and np.isnan(r[1]) and r[2] == inf
    r = po.math("exp", [1000.0, -1000.0, np.nan])
    assert r[0] == inf and r[1] == 0.0 and np.isnan(r[2])
    r = po.math("log1pf", [-1.0, -2.0, inf, 0.0])
    assert r[0] == -inf and np.isnan(r[1]) and r[2] == inf and r[3] == 0.0
    r = po.math("softplus", [-800.0, 0.0, 800.0])
    assert r[0] == 0.0 and abs(r[1] - np.log(2)) < 1e-16 and r[2] == 800.0


def test_fixed_point_round_trip_and_clamp():
    v = np.array([0.0, 1.0, -1.0, 2.0 ** -32, 3.0 * 2.0 ** -33, 123456.789, -1e9, np.nan, 1e30, -1e30])
    r = po.math("fix", v)
    assert r[0] == 0.0 and r[1] == 1.0 and r[2] == -1.0 and r[3] == 2.0 ** -32
    assert r[4] == 2.0 ** -31          # 1.5 ulp rounds to even
    assert abs(r[5] - 123456.789) <= 2.0 ** -33
    assert r[6] == -1e9
    assert r[7] == -4.6e18 * 2.0 ** -32   # NaN maps to the negative clamp
    assert r[8] == 4.6e18 * 2.0 ** -32 and r[9] == -4.6e18 * 2.0 ** -32


# --- samplers -------------------------------------------------------------------

def draws(kind, n, p1=0.0, p2=0.0, seed=4, phase=10):
    return po.stream(kind, n, seed=seed, phase=phase, p1=p1, p2=p2)


def test_normal_distribution():
    x = draws("normal", 100000)
    assert st.kstest(x, "norm").pvalue > 1e-3
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1) < 0.01


@pytest.mark.parametrize("a,scale", [(0.05, 1.0), (0.3, 2.0), (1.0, 1.0), (3.7, 0.5), (50.0, 0.1)])
def test_gamma_distribution(a, scale):
    x = draws("gamma", 40000, a, scale)
    assert (x >= 0).all()
    assert st.kstest(x, st.gamma(a, scale=scale).cdf).pvalue > 1e-3


@pytest.mark.parametrize("lam", [0.3, 4.0, 9.99, 10.0, 47.0, 3000.0])
def test_poisson_distribution(lam):
    x = draws("poisson", 40000, lam).astype(int)
    lo, hi = st.poisson.ppf(1e-4, lam), st.poisson.ppf(1 - 1e-4, lam)
    edges = np.unique(np.linspace(lo, hi + 1, 30).astype(int))
    obs = np.histogram(np.clip(x, lo, hi), bins=edges)[0]
    p = np.diff(st.poisson.cdf(edges - 1, lam))
    p[0] += st.poisson.cdf(lo - 1, lam)
    p[-1] += st.poisson.sf(hi, lam)
    exp = p / p.sum() * len(x)
    keep = exp > 5
    chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert st.chi2.sf(chi, keep.sum() - 1) > 1e-3
    assert abs(x.mean() - lam) < 5 * np.sqrt(lam / len(x))


NB_CASES = [(0.5, 5.0), (20.0, 5.0), (29.9, 5.3), (300.0, 0.7), (5.0, 0.05), (31.0, 50.0)]


@pytest.mark.parametrize("kind", ["nbinom", "nbinom_xz"])
@pytest.mark.parametrize("mu,phi", NB_CASES)
def test_negative_binomial_distribution(kind, mu, phi):
    x = draws(kind, 40000, mu, phi).astype(int)
    dist = st.nbinom(phi, phi / (phi + mu))
    assert abs(x.mean() - mu) < 6 * np.sqrt(dist.var() / len(x))
    hi = int(dist.ppf(0.999))
    obs = np.bincount(np.minimum(x, hi), minlength=hi + 1)
    p = dist.pmf(np.arange(hi + 1))
    p[-1] += dist.sf(hi)
    exp = p * len(x)
    keep = exp > 5
    chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert st.chi2.sf(chi, keep.sum() - 1) > 1e-3


def test_negative_binomial_inversion_is_the_exact_quantile():
    # bq_rnbinom_inv(u) must be the smallest k with CDF(k) >= u (scipy's ppf) over the
    # admissible region mu <= 30, mu/(phi+mu) <= 0.85, including u next to 1
    rng = np.random.default_rng(5)
    n = 3000
    mu = np.exp(rng.uniform(np.log(1e-3), np.log(30.0), n))
    phi = mu * (0.15 / 0.85) * (1 + 1e-9) * np.exp(rng.uniform(0, np.log(1e4), n))
    u = (rng.integers(0, 2 ** 32, n) + 0.5) * 2.0 ** -32
    u[:20] = 1 - 2.0 ** -33
    assert ((mu <= 30) & (3 * mu <= 17 * phi)).all()
    got = po.nb_inv(u, mu, phi)
    ref = st.nbinom.ppf(u, phi, phi / (phi + mu))
    assert got.max() <= 255
    # a draw may differ only where u lies within rounding of a CDF step
    bad = got != ref
    step = np.abs(st.nbinom.cdf(np.minimum(got, ref), phi, phi / (phi + mu)) - u)
    assert bad.sum() <= 2 and (step[bad] < 1e-12).all(), (got[bad], ref[bad])
    # planted mistake: a mean off by 5% is caught
    assert (po.nb_inv(u, mu * 1.05, phi) != ref).sum() > 50


@pytest.mark.parametrize("a,b", [(2.0, 5.0), (1.0, 3.0), (40.0, 3000.0)])
def test_beta_distribution(a, b):
    x = draws("beta", 30000, a, b)
    assert st.kstest(x, st.beta(a, b).cdf).pvalue > 1e-3


def test_planted_mistake_is_detected():
    # a shifted normal (a plausible sign/offset slip) must fail the same KS test
    x = draws("normal", 100000) + 0.02
    assert st.kstest(x, "norm").pvalue < 1e-3


def test_single_precision_ratio_is_within_two_roundings_and_never_above_one():
    # r = mu (1 / (phi + mu)) in single precision: within 2^-22 of the exact
    # mu / (phi + mu) of the float inputs, and held to at most 1, which the
    # subtype term needs (r expm1(beta) > -1 for expm1(beta) >= -1 + 2^-24)
    rng = np.random.default_rng(17)
    n = 200000
    mu = np.exp(rng.uniform(np.log(1e-30), np.log(1e30), n)).astype(np.float32)
    phi = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n)).astype(np.float32)
    phi[:2000] = mu[:2000] * np.float32(1e-9)   # phi negligible: phi + mu == mu, r rounds to 1 or just below
    phi[2000:2100] = np.float32(1e-38)
    mu[2100:2200] = np.float32(3e38)            # phi + mu beyond the reciprocal's hold at 2^125
    r = po.terms("ratio", 32, np.stack([mu.astype(np.float64), phi.astype(np.float64)], 1))
    exact = mu.astype(np.float64) / (phi.astype(np.float64) + mu.astype(np.float64))
    assert r.max() <= 1.0
    ok = mu < 2.0 ** 124
    assert np.all(np.abs(r[ok] - exact[ok]) <= 2.0 ** -22 * exact[ok] + 1e-45)
    assert np.all(r[~ok] == 1.0)
    # the subtype term stays finite at the smallest held expm1(beta) with r = 1
    em = np.full(2000, -1 + 2.0 ** -24)
    tw = po.terms("tw", 32, np.stack([np.full(2000, 3.0), phi[:2000].astype(np.float64), r[:2000],
                                      np.log1p(em), em], 1))
    assert np.isfinite(tw).all()
    # planted mistake: without the minimum, mu beyond the reciprocal's hold gives r > 1
    unheld = mu[2100:2200] * (np.float32(1) / np.float32(2.0 ** 125))
    assert unheld.min() > 1.0


def test_random_walk_log1p_arguments_stay_in_the_branch_free_range():
    # The single-precision MH terms take log1p(r c) without a range branch: r in
    # [0, 1], c = expm1 of one or two steps of 0.1 times a normal deviate, and
    # that deviate inverts a uniform in [2^-33, 1 - 2^-33]
    z = po.math("ndtri", [2.0 ** -33, 1 - 2.0 ** -33])
    assert abs(z[0]) <= 6.35 and abs(z[1]) <= 6.35 and z[0] == -z[1]
    step = 0.1 * np.abs(z).max()
    c = np.array([np.expm1(-step), np.expm1(step), np.expm1(-2 * step), np.expm1(2 * step)])
    assert c.min() > -0.73 and c.max() < 2.6
    # over that range (and r in [0, 1]) the branch-free log1pf is within 2 ulp of mpmath
    import mpmath as mp
    y = np.float32(np.linspace(c.min(), c.max(), 20001))
    ref = np.array([float(mp.log1p(mp.mpf(float(v)))) for v in y])
    assert ulps(po.math("log1pf_mid", y.astype(np.float64)), ref, np.float32).max() <= 2.0
    # the MH term with r = 1 and the extreme steps is finite
    t = po.terms("tmh", 32, np.array([[5.0, 2.0, 1.0, -2 * step, c[2]], [5.0, 2.0, 1.0, 2 * step, c[3]]]))
    assert np.isfinite(t).all()


HOST_PASS_SRC = r"""
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "busseq/numerics.h"
__global__ void k_uses_tables(double* o) { o[threadIdx.x] = bq_log(2.0 + threadIdx.x) + bq_exp(0.5) + bq_inv_tab[3]; }
int main(int argc, char** argv) {
    for (int j = 1; j < argc; ++j) {
        double x;
        sscanf(argv[j], "%la", &x);
        printf("%a %a %a %a\n", bq_log(x), bq_exp(x - 1.0), bq_log1p(x - 1.0), bq_lgamma(x));
    }
    printf("%a\n", bq_inv_tab[3]);
    return 0;
}
"""


def test_host_pass_of_a_hip_source_reads_the_constant_tables(tmp_path):
    # Host code compiled by hipcc (the host side of engine.hip, init.hip) evaluates the
    # same numerics as the gcc-built oracle: the coefficient and reciprocal tables that
    # the device reads from constant memory are ordinary initialised arrays there.
    import subprocess
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc is not installed")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "host_pass.hip"
    src.write_text(HOST_PASS_SRC)
    exe = tmp_path / "host_pass"
    subprocess.run([hipcc, "-O2", "--offload-arch=gfx950", "-ffp-contract=off", "-fno-fast-math",
                    "-I", inc, str(src), "-o", str(exe)], check=True, capture_output=True)
    xs = np.array([0.3, 1.0, 2.0, 7.25, 123.5, 1e-3, 4096.0])
    out = subprocess.run([str(exe)] + [float(x).hex() for x in xs], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    got = np.array([[float.fromhex(v) for v in line.split()] for line in out[:len(xs)]])
    assert np.array_equal(got[:, 0], po.math("log", xs))
    assert np.array_equal(got[:, 1], po.math("exp", xs - 1.0))
    assert np.array_equal(got[:, 2], po.math("log1p", xs - 1.0))
    assert np.array_equal(got[:, 3], po.math("lgamma", xs))
    assert float.fromhex(out[len(xs)]) == 1.0 / 3.0
