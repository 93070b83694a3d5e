"""Write tests/golden/math_vectors.npz: inputs and 40-digit mpmath values of
log, log1p, exp, expm1, lgamma, lgamma differences and the standard-normal
quantile, rounded to double.  Run: python tests/golden/make_math_vectors.py
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
rng = np.random.default_rng(20261203)
out = {}


def vec(fn, xs):
    return np.array([float(fn(mp.mpf(float(x)))) for x in xs])


x = np.concatenate([np.exp(rng.uniform(-700, 700, 1500)), [1.0, 2.0, 0.5, 1e-310, 4.9e-324, 1.7e308]])
out["log_x"], out["log_y"] = x, vec(mp.log, x)
x = np.concatenate([rng.uniform(-0.999999, 3, 1500), rng.uniform(-1e-6, 1e-6, 200),
                    np.exp(rng.uniform(-40, 40, 300)), [0.0, -0.5, 0.41421356237309515, -0.29289321881345248]])
out["log1p_x"], out["log1p_y"] = x, vec(mp.log1p, x)
x = np.concatenate([rng.uniform(-745, 709.7, 1500), rng.uniform(-1, 1, 300), [0.0, -740.0, 709.0]])
out["exp_x"], out["exp_y"] = x, vec(mp.exp, x)
x = np.concatenate([rng.uniform(-50, 50, 800), rng.uniform(-0.4, 0.4, 800), [0.0, 1e-300, -1e-20]])
out["expm1_x"], out["expm1_y"] = x, vec(mp.expm1, x)
x = np.concatenate([np.exp(rng.uniform(-12, 14, 1500)), [1.0, 2.0, 0.5, 3.0, 10.0, 9.999999]])
out["lgamma_x"], out["lgamma_y"] = x, vec(mp.loggamma, x)
z = np.exp(rng.uniform(-3, 9, 1500))
d = rng.uniform(-0.95, 2.0, 1500) * np.minimum(z, 8.0)
out["lgdiff_z"], out["lgdiff_d"] = z, d
out["lgdiff_y"] = np.array([float(mp.loggamma(mp.mpf(float(a)) + mp.mpf(float(b))) - mp.loggamma(mp.mpf(float(a))))
                            for a, b in zip(z, d)])
p = np.concatenate([rng.uniform(0, 1, 1500), 10.0 ** rng.uniform(-300, -1, 200), [0.5, 0.025, 0.975]])
out["ndtri_x"] = p


def ndtri(v):
    with mp.workdps(360):   # 2p - 1 must keep the tiny tail probabilities
        return float(mp.sqrt(2) * mp.erfinv(2 * mp.mpf(float(v)) - 1))


out["ndtri_y"] = np.array([ndtri(v) for v in p])
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "math_vectors.npz"), **out)
print({k: v.shape for k, v in out.items()})
