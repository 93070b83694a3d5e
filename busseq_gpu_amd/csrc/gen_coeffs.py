"""Fit the polynomial coefficients used by include/busseq/numerics.h.

Chebyshev interpolation (near-minimax) of smooth kernels, done in 60-digit
mpmath arithmetic, converted to the monomial basis and rounded to the target
precision.  Prints C initialisers.  Run: python busseq_gpu_amd/csrc/gen_coeffs.py
"""
import mpmath as mp

mp.mp.dps = 60


def cheb_fit(fn, a, b, deg):
    a, b = mp.mpf(a), mp.mpf(b)
    n = deg + 1
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    xs = [(b - a) / 2 * t + (a + b) / 2 for t in nodes]
    fx = [fn(x) for x in xs]
    # Chebyshev coefficients on [-1,1]
    c = []
    for j in range(n):
        s = mp.fsum(fx[k] * mp.cos(j * mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n))
        c.append(2 * s / n)
    c[0] /= 2
    # convert sum c_j T_j(t) with t=(2x-a-b)/(b-a) to monomials in x
    # build T_j as polynomials in t, then substitute
    T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for j in range(2, n):
        p = [mp.mpf(0)] + [2 * v for v in T[j - 1]]
        for i, v in enumerate(T[j - 2]):
            p[i] -= v
        T.append(p)
    poly_t = [mp.mpf(0)] * n
    for j in range(n):
        for i, v in enumerate(T[j]):
            poly_t[i] += c[j] * v
    # t = alpha*x + beta
    alpha = 2 / (b - a)
    beta = -(a + b) / (b - a)
    poly_x = [mp.mpf(0)] * n
    # expand (alpha x + beta)^i
    for i, ci in enumerate(poly_t):
        for m in range(i + 1):
            poly_x[m] += ci * mp.binomial(i, m) * alpha ** m * beta ** (i - m)
    return poly_x


def horner(coefs, x):
    r = mp.mpf(0)
    for cf in reversed(coefs):
        r = r * x + cf
    return r


def report(name, fn, a, b, deg, cast, rel_fn):
    co = cheb_fit(fn, a, b, deg)
    co_r = [mp.mpf(cast(float(v))) for v in co]
    worst = mp.mpf(0)
    for k in range(4001):
        x = mp.mpf(a) + (mp.mpf(b) - mp.mpf(a)) * k / 4000
        if x == 0:
            continue
        err = rel_fn(x, horner(co_r, x))
        worst = max(worst, abs(err))
    return co_r, worst


def log1p_q(x):
    # Q(f) = (log1p(f) - f) / f^2
    if x == 0:
        return mp.mpf(-0.5)
    return (mp.log1p(x) - x) / x ** 2


def exp_e(x):
    # E(r) = (exp(r) - 1 - r) / r^2
    if x == 0:
        return mp.mpf(0.5)
    return (mp.expm1(x) - x) / x ** 2


def stirling_q(u):
    # Q(u) = w S(w) with u = 1/w^2 and S(w) = lgamma(w) - (w - 1/2) log w + w - log(2 pi)/2
    if u == 0:
        return mp.mpf(1) / 12
    w = 1 / mp.sqrt(u)
    return w * (mp.loggamma(w) - (w - mp.mpf(1) / 2) * mp.log(w) + w - mp.log(2 * mp.pi) / 2)


def main():
    import numpy as np
    f32 = lambda v: float(np.float32(v))
    f64 = lambda v: float(v)
    lo, hi = mp.sqrt(0.5) - 1, mp.sqrt(2) - 1
    lo = lo - mp.mpf("1e-3")
    hi = hi + mp.mpf("1e-3")

    def rel_log(x, q):
        return (x + x * x * q - mp.log1p(x)) / mp.log1p(x)

    def rel_exp(x, e):
        return (1 + x + x * x * e - mp.exp(x)) / mp.exp(x)

    out = []
    for name, cast, deg in (("LOG1P_Q_F32", f32, 8), ("LOG1P_Q_F32_MID", f32, 7), ("LOG1P_Q_F64", f64, 19)):
        co, w = report(name, log1p_q, lo, hi, deg, cast, rel_log)
        out.append((name, co, w))
    r = mp.log(2) / 2 + mp.mpf("1e-3")
    for name, cast, deg in (("EXP_E_F64", f64, 11),):
        co, w = report(name, exp_e, -r, r, deg, cast, rel_exp)
        out.append((name, co, w))
    # Stirling remainder for w >= 1 (u in (0, 1]); the error is absolute, in S = Q(u)/w
    def abs_s(u, q):
        return (q - stirling_q(u)) * mp.sqrt(u)

    co, w = report("STIRLING_Q_F32", stirling_q, 0, 1, 6, f32, abs_s)
    out.append(("STIRLING_Q_F32", co, w))

    # double precision, 1 <= w <= 10: w S(w) in t = 1/w on [0.1, 1] (slightly
    # widened); w >= 10 keeps the Stirling series.  The error is absolute, in S = t Q(t)
    def stirling_t(t):
        return stirling_q(t * t)

    def abs_st(t, q):
        return (q - stirling_t(t)) * t

    co, w = report("STIRLING_T_F64", stirling_t, mp.mpf("0.0999"), mp.mpf("1.0001"), 20, f64, abs_st)
    out.append(("STIRLING_T_F64", co, w))
    for name, co, w in out:
        print(f"/* {name}: degree {len(co)-1}, max err of the rounded polynomial {mp.nstr(w, 3)} */")
        print(f"#define {name} \\")
        print("    " + ", \\\n    ".join(mp.nstr(v, 21, min_fixed=0, max_fixed=0) for v in co))
        print()


if __name__ == "__main__":
    main()
