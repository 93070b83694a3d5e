"""Error of the mixed-precision sweep against the all-FP64 oracle, per path,
on the 16 heterogeneous data sets (N 500-30000, G 1000-5000, B 2-6, K 3-10;
busseq_gpu_amd/workloads.py).

Each data set is burnt in for 40 sweeps by the sampler itself (FP64 terms, on
the GPU); from that state an FP32 and an FP64 oracle run the same sweep on the
same Philox streams, and the quantity each path decides on is compared: the
subtype logits (sum over genes), the alpha, beta, nu sums (over a batch's or a
subtype's cells), the delta sums (over genes), the phi sums (over a batch's
cells, lgamma differences) and the gamma sums (the X histogram, FP64 in both).
Run with -s for the table.  Where the two precisions drew a different X for
an entry (the FP32 mean moves an NB draw across a CDF step) or decided a
parameter differently, every later sum of that sweep moves by more than
rounding; such a data set contributes only its subtype logits (computed first,
from the identical state), and the number of them is reported and bounded.
"""
import numpy as np
import pytest

import busseq_gpu_amd as bq
from busseq_gpu_amd import workloads as wl
from oracle import pyoracle as po
from parity import oracle_threads

pytestmark = pytest.mark.gpu

PATHS = ("w_logit", "alpha_sum", "beta_sum", "nu_sum", "delta_sum", "phi_sum", "gamma_sum")
# bounds of |FP32 - FP64| <= abs + rel |FP64 value|, from the measured values with
# a margin: the sums over cells grow with the cells summed (up to 21300 here),
# because each gene's single-precision parameter copy (e^alpha, e^beta, e^nu, phi)
# carries one rounding that every cell of the gene shares; the phi sums' largest
# differences come with proposals phi' < 0.03, whose log-ratios (-300 to -35000)
# are rejected whatever the rounding, so they are bounded relative to the value
BOUNDS = {"w_logit": (1e-2, 0), "alpha_sum": (5e-2, 0), "beta_sum": (5e-2, 0), "nu_sum": (5e-2, 0),
          "delta_sum": (1e-2, 0), "phi_sum": (0.1, 1e-2), "gamma_sum": (1e-9, 0)}


@pytest.mark.timeout(1200)
def test_mixed_precision_error_per_path_on_sixteen_data_sets():
    phase = dict(n_iter=400, n_burnin=200, n_unchanged=20)
    probs = wl.c5_problems(bq, **phase)
    print(f"\n[precision] {len(probs)} data sets, burn-in on the GPU (FP64)", flush=True)
    with bq.Sampler.batch(probs, precision=64) as s:
        s.sweeps(40)
        states = [s.state(j) for j in range(len(probs))]
    errs = {k: [] for k in PATHS}
    vals = {k: [] for k in PATHS}
    moved, entries, decisions, diverged, n_dec = 0, 0, 0, 0, 0
    for j, (p, st) in enumerate(zip(probs, states)):
        out = {}
        for prec in (32, 64):
            o = po.Oracle(p["Y"], p["n_b"], p["K"], seed=31 + j, precision=prec, threads=oracle_threads(), **phase)
            o.set_state(st)
            o.sweeps(1)
            out[prec] = (o.debug(), o.state())
            o.close()
        (d32, s32), (d64, s64) = out[32], out[64]
        N, G, B, K = p["Y"].shape[0], p["Y"].shape[1], len(p["n_b"]), p["K"]
        nx = int(np.sum(s32["S"] != s64["S"]))
        nd = sum(int(np.sum(s32[k] != s64[k])) for k in ("W", "alpha", "beta", "L", "nu", "delta", "phi", "gamma", "pi"))
        moved += nx
        entries += s32["S"].size
        decisions += nd
        n_dec += int(np.sum(p["Y"] == 0)) + N + G + 2 * G * (K - 1) + G * (B - 1) + N + B * G + 2 * B + B * K
        # the subtype logits come first in the sweep, from the identical state; the
        # other sums follow the X/Z draws and the decisions before them
        clean = nx == 0 and nd == 0
        diverged += not clean
        for k in PATHS if clean else ("w_logit",):
            errs[k].append(np.abs(np.asarray(d32[k]) - np.asarray(d64[k])).ravel())
            vals[k].append(np.abs(np.asarray(d64[k])).ravel())
        print(f"[precision] data set {j}: N={p['Y'].shape[0]} G={p['Y'].shape[1]} B={len(p['n_b'])} K={p['K']}: "
              f"{nx} entries with a different X/Z draw, {nd} different decisions", flush=True)
    print("\npath        items   median |v|   max |d|   max |d|/max(1,|v|)   99.9% |d|   median |d|"
          "   max |d| where |v| <= 100")
    for k in PATHS:
        e, v = np.concatenate(errs[k]), np.concatenate(vals[k])
        small = v <= 100
        print(f"{k:10s} {e.size:8d} {np.median(v):12.4g} {e.max():9.3g} {np.max(e / np.maximum(1, v)):20.3g}"
              f" {np.quantile(e, 0.999):11.3g} {np.median(e):11.3g} {e[small].max() if small.any() else 0:14.3g}")
        ab, rel = BOUNDS[k]
        assert np.all(e <= ab + rel * v), (k, e.max())
    print(f"entries whose X or Z differs between the precisions: {moved} of {entries}; "
          f"parameter or subtype values that differ after the sweep: {decisions} (of {n_dec} draws and "
          f"decisions); data sets whose sweep diverged: {diverged} of {len(probs)}")
    assert moved + decisions <= 1e-4 * n_dec
