"""Error of the mixed-precision sweep against the all-FP64 oracle, per path.

Mixed precision (precision 32) evaluates every per-element term -- the NB
log-pmf differences of the alpha, beta, nu, delta and phi MH ratios
(BUSseq_gpu.cu:608-632, 676-818), their lgamma differences, and the subtype
logits that replace the W step of cu:820-854 -- in FP32, and adds the terms
exactly in fixed point (FP64 partials per group); the gamma sums
(cu:524-600) are built from the X histogram and the BIC (cu:967-1024) in FP64
in both modes.  From one state reached by an FP64 chain, an FP32 and an FP64
oracle run the same sweep on the same Philox streams; the quantity each path
decides on is compared.  Run with -s to print the table of errors.

The bounds are those measured on this workload with a margin: a few 1e-4 in
log units for the sums over G genes or over a batch's cells, about 1e-2 for
the phi sums (lgamma differences at large counts carry most of it), and zero
for gamma.  No accept/reject decision differs in the sweep.
"""
import numpy as np
import pytest

import busseq_gpu_amd as bq
from oracle import pyoracle as po

BOUNDS = {"w_logit": 2e-3, "alpha_sum": 2e-3, "beta_sum": 2e-3, "nu_sum": 2e-3, "delta_sum": 2e-3,
          "phi_sum": 5e-2, "gamma_sum": 0.0}


@pytest.fixture(scope="module")
def one_sweep_both_precisions():
    n_b, G, K = [300, 250, 200, 150], 1000, 5
    Y = bq.simulate(n_b, G, K, 1005)
    burn = po.Oracle(Y, n_b, K, seed=777, precision=64, n_iter=400, n_burnin=200, n_unchanged=50)
    burn.sweeps(40)
    start = burn.state()
    out = {}
    for prec in (32, 64):
        o = po.Oracle(Y, n_b, K, seed=31, precision=prec, n_iter=400, n_burnin=200, n_unchanged=50)
        o.set_state(start)
        o.sweeps(1)
        out[prec] = (o.debug(), o.state())
    return out


def test_mixed_precision_error_per_path(one_sweep_both_precisions):
    d32, _ = one_sweep_both_precisions[32]
    d64, _ = one_sweep_both_precisions[64]
    print("\npath        median |FP64 value|   max |FP32 - FP64|   99.9% |FP32 - FP64|   median")
    for key, bound in BOUNDS.items():
        err = np.abs(d32[key] - d64[key])
        print(f"{key:10s} {np.median(np.abs(d64[key])):18.4g} {err.max():18.3g} {np.quantile(err, 0.999):20.3g}"
              f" {np.median(err):10.3g}")
        assert err.max() <= bound, (key, err.max())


def test_mixed_precision_makes_the_same_decisions(one_sweep_both_precisions):
    _, s32 = one_sweep_both_precisions[32]
    _, s64 = one_sweep_both_precisions[64]
    for key in ("S", "W", "alpha", "beta", "L", "nu", "delta", "phi", "gamma", "pi"):
        assert np.array_equal(s32[key], s64[key]), key


def test_the_comparison_detects_a_single_precision_slip(one_sweep_both_precisions):
    # planted mistake: the phi sums of the FP32 run shifted by one FP32 ulp of their
    # largest term would still pass; shifted by 0.1 (a dropped lgamma term is far larger) they fail
    d32, _ = one_sweep_both_precisions[32]
    d64, _ = one_sweep_both_precisions[64]
    assert np.abs((d32["phi_sum"] + 0.1) - d64["phi_sum"]).max() > BOUNDS["phi_sum"]
