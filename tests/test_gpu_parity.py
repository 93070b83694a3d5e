"""GPU parity: the gfx950 sampler against the CPU oracle, sweep by sweep.

Both are driven by the same Philox streams and the same numerics
(include/busseq/numerics.h, model.h) and use the same reduction grouping, so
integer draws (X, Z, W, L) must match bit for bit and the continuous
parameters must be identical (rtol 1e-12 written below; equality is expected).
"""
import numpy as np
import pytest

import busseq_gpu_amd as bq
from oracle import pyoracle as po
from parity import compare

pytestmark = pytest.mark.gpu


def run_pair(Y, n_b, K, steps, prec=32, tile_rows=64, **extra):
    kw = dict(seed=extra.pop("seed", 13579), n_iter=extra.pop("n_iter", 40), n_burnin=extra.pop("n_burnin", 10),
              n_unchanged=extra.pop("n_unchanged", 5), precision=prec)
    print(f"[parity] prec={prec} N={sum(n_b)} G={Y.shape[1]} K={K} B={len(n_b)} TR={tile_rows}", flush=True)
    o = po.Oracle(Y, n_b, K, tile_rows=tile_rows, **kw)
    s = bq.Sampler(Y, n_b, K, tile_rows=tile_rows, **kw)
    compare(s.state(), o.state(), "init")
    for step in steps:
        s.sweeps(step)
        o.sweeps(step)
        compare(s.state(), o.state(), f"iter {o.iteration}")
    return s, o


@pytest.mark.parametrize("prec", [32, 64])
def test_sweeps_match_oracle(small_data, prec):
    d = small_data
    s, o = run_pair(d["Y"], d["n_b"], d["K"], (1, 1, 3, 10), prec)
    s.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_posterior_and_bic_match_oracle(small_data, prec):
    d = small_data
    s, o = run_pair(d["Y"], d["n_b"], d["K"], (40,), prec, n_iter=40, n_burnin=25, n_unchanged=5)
    r = s.finish()
    p = o.posterior()
    for k in ("gamma", "alpha", "L", "beta", "ptau", "nu", "delta", "phi", "pi"):
        np.testing.assert_allclose(r[k], p[k], rtol=1e-12, atol=1e-300, err_msg=k)
    np.testing.assert_array_equal(r["W"], p["W"])
    np.testing.assert_allclose(r["bic"], p["bic"], rtol=1e-12)
    s.close()


def test_bic_entry_point_equals_the_oracle_at_any_parameters(small_data):
    # busseq_bic at the posterior means is finish()'s BIC; at another parameter
    # set (the chain's current state) it equals the oracle's BIC there
    d = small_data
    s, o = run_pair(d["Y"], d["n_b"], d["K"], (30,), 32, n_iter=30, n_burnin=20, n_unchanged=5)
    r = s.finish()
    assert s.bic(r) == r["bic"]
    st = s.state()
    np.testing.assert_allclose(s.bic(st), o.bic(st), rtol=1e-12)
    np.testing.assert_allclose(s.bic(o.posterior()), o.posterior()["bic"], rtol=1e-12)
    s.close()


def test_tile_rows_and_k2_b2():
    n_b = [9, 13]
    Y = bq.simulate(n_b, 70, 2, 31)
    s, o = run_pair(Y, n_b, 2, (1, 4, 5), 32, tile_rows=4)
    s.close()


def test_large_counts_overflow_histogram():
    # counts >= 1024 take the direct gamma-term path instead of the histogram
    n_b = [20, 20]
    Y = bq.simulate(n_b, 96, 3, 5)
    Y[::3, ::7] = Y[::3, ::7] * 400 + 1500
    s, o = run_pair(Y, n_b, 3, (1, 2, 4), 32)
    s.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_counts_at_the_phi_queue_bound(prec):
    # the phi step takes the lgamma parts of counts 1..127 from the static
    # histogram's list, queues the others below 2^26 and adds the larger ones
    # inline: 127 and 128, 2^26 - 1, 2^26 and 7e7 take the three paths
    n_b = [20, 15]
    Y = bq.simulate(n_b, 40, 3, 23)
    Y[1, 5], Y[7, 11], Y[12, 3], Y[30, 5] = 2**26 - 1, 2**26, 70_000_000, 2**26
    Y[2:9, 7], Y[9:14, 7], Y[21:30, 8] = 127, 128, 127
    s, o = run_pair(Y, n_b, 3, (1, 2, 3), prec)
    s.close()


def test_edge_batch_equals_k_and_zero_gene():
    n_b = [4, 4, 4]
    Y = bq.simulate(n_b, 65, 4, 11)
    Y[:, 3] = 0    # a gene that is never observed
    s, o = run_pair(Y, n_b, 4, (1, 5), 64)
    s.close()


def test_batched_chains_equal_plain_chains(small_data):
    d = small_data
    kw = dict(seed=99, n_iter=30, n_burnin=10, n_unchanged=5, precision=32)
    with bq.Sampler(d["Y"], d["n_b"], d["K"], chains=3, **kw) as sb:
        sb.sweeps(12)
        for c in range(3):
            with bq.Sampler(d["Y"], d["n_b"], d["K"], chains=1, chain0=c, **kw) as sp:
                sp.sweeps(12)
                compare(sb.state(c), sp.state(), f"chain {c}")


def test_set_state_roundtrip(small_data):
    d = small_data
    kw = dict(seed=5, n_iter=30, n_burnin=10, n_unchanged=5, precision=32)
    o = po.Oracle(d["Y"], d["n_b"], d["K"], **kw)
    o.sweeps(3)
    st = o.state()
    with bq.Sampler(d["Y"], d["n_b"], d["K"], **kw) as s:
        s.sweeps(3)
        compare(s.state(), st, "iter 3")
        # perturb, then restore the oracle state and continue both
        s.set_state({"alpha": st["alpha"] + 0.5})
        s.set_state({k: st[k] for k in ("X", "Z", "W", "alpha", "beta", "L", "nu", "delta", "phi",
                                         "gamma", "pi", "ptau")})
        s.sweeps(2)
        o.sweeps(2)
        compare(s.state(), o.state(), "after set_state")
        # a state neither side reached by sweeping: the proposals and tables the
        # next sweep uses must come from it, not from the last sweep's parameters
        # (scalings keep the parameters fixed at zero there: beta of subtype 0,
        # nu of batch 0, delta of each batch's first cell)
        st = o.state()
        nu = st["nu"].copy()
        nu[1:] += 0.0625
        moved = {"alpha": st["alpha"] - 0.25, "delta": st["delta"] * 0.875, "gamma": st["gamma"] * 0.75,
                 "nu": nu, "beta": st["beta"] * 0.5}
        s.set_state(moved)
        o.set_state(moved)
        s.sweeps(2)
        o.sweeps(2)
        compare(s.state(), o.state(), "after setting a new state")


@pytest.mark.slow
def test_c1_shape_parity():
    n_b = [300, 300, 200, 200]
    Y = bq.simulate(n_b, 3000, 5, 1001)
    s, o = run_pair(Y, n_b, 5, (1, 2), 32, n_iter=4000, n_burnin=2000, n_unchanged=500)
    s.close()


def test_cell_without_counts_is_rejected():
    # the device init pass reports a zero library size (the reference would divide by it)
    n_b = [6, 6]
    Y = bq.simulate(n_b, 40, 2, 3)
    Y[7, :] = 0
    with pytest.raises(bq.BusseqError) as e:
        bq.Sampler(Y, n_b, 2, n_iter=4, n_burnin=2, n_unchanged=1)
    assert e.value.status == 4 and "cell 7" in str(e.value)


def test_raw_means_span_batches_and_gene_slabs():
    # several 64-cell chunks per batch, batch ends inside a chunk, G not a multiple of 256
    n_b = [150, 70, 131]
    Y = bq.simulate(n_b, 300, 3, 17)
    s, o = run_pair(Y, n_b, 3, (1,), 32, tile_rows=16)
    s.close()


def test_graph_equals_in_order_launches(small_data):
    # the captured sweep and the per-kernel timing path (events between the
    # launches) must both give the oracle's state bit for bit
    d = small_data
    kw = dict(seed=23, n_iter=30, n_burnin=4, n_unchanged=2, precision=32)
    o = po.Oracle(d["Y"], d["n_b"], d["K"], **kw)
    o.sweeps(6)
    with bq.Sampler(d["Y"], d["n_b"], d["K"], **kw) as sg, bq.Sampler(d["Y"], d["n_b"], d["K"], **kw) as ss:
        sg.sweeps(6)
        ss.kernel_times(6)
        assert ss.iteration == sg.iteration == 6
        compare(sg.state(), o.state(), "graph")
        compare(ss.state(), o.state(), "in order")


@pytest.mark.parametrize("n_b, G, K, prec", [
    ([30, 24, 20, 16], 130, 7, 32),                          # the subtype count of N=20000, G=3000, K=7
    ([30, 24, 20, 16], 130, 7, 64),
    ([26, 22, 20, 18, 16, 15, 14, 13], 200, 12, 32),         # B=8, K=12 as at N=150000, G=5000
    ([20, 17, 17], 70, 16, 32),                              # the largest K with a kernel specialised on K
    ([30, 28, 26], 150, 20, 32),                             # K > 16: subtype groups of 16 and 3
    ([40, 36], 120, 33, 32),                                 # two full groups of 16
    ([50, 46], 100, 40, 64),                                 # three groups, FP64
    ([70, 66], 64, 64, 32),                                  # the largest K: four groups, lane k = subtype k
])
def test_large_k_instantiations_match_oracle(n_b, G, K, prec):
    Y = bq.simulate(n_b, G, K, 40 + K)
    s, o = run_pair(Y, n_b, K, (1, 2, 3), prec, tile_rows=8, n_iter=20, n_burnin=4, n_unchanged=2)
    s.close()
