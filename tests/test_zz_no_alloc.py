"""busseq_bic_noalloc, the allocation-free form of busseq_bic, through the C ABI.

Two calls in a row on different chains, the second on a smaller problem (the
shared device array keeps the first call's contents), each at two parameter
sets: every result equals busseq_bic's bit for bit.  A chain beyond the
array's stated size is rejected with BUSSEQ_ERR_ARG and a message.
"""
import ctypes as C

import numpy as np
import pytest

import busseq_gpu_amd as bq
from busseq_gpu_amd import _dp, _load

pytestmark = pytest.mark.gpu


def noalloc_bic(s, params, chain=0):
    lib = _load()
    arrs = [np.ascontiguousarray(params[k], np.float64) for k in bq.BIC_KEYS]
    out = C.c_double(np.nan)
    st = lib.busseq_bic_noalloc(s._h, chain, *[_dp(a) for a in arrs], C.byref(out))
    return st, out.value


def test_noalloc_bic_equals_busseq_bic_on_two_calls_the_second_smaller():
    problems = [([60, 50, 40], 300, 4, 7), ([20, 15], 100, 3, 8)]
    for n_b, G, K, ds in problems:
        Y = bq.simulate(n_b, G, K, ds)
        print(f"[noalloc] N={sum(n_b)} G={G} K={K}", flush=True)
        with bq.Sampler(Y, n_b, K, seed=ds, n_iter=20, n_burnin=10, n_unchanged=3) as s:
            s.sweeps(20)
            r = s.finish()
            for params in (r, s.state()):
                st, v = noalloc_bic(s, params)
                assert st == 0, bq._load().busseq_last_error().decode()
                assert v == s.bic(params)
            assert noalloc_bic(s, r)[1] == r["bic"]


def test_noalloc_bic_rejects_a_chain_beyond_its_scratch():
    n_b, G, K = [135000, 135000], 64, 16      # N K = 4.32M doubles > 4194304
    Y = bq.simulate(n_b, G, K, 9)
    print(f"[noalloc] N={sum(n_b)} G={G} K={K} (too large)", flush=True)
    with bq.Sampler(Y, n_b, K, seed=3, n_iter=4, n_burnin=2, n_unchanged=1) as s:
        st = s.state()
        params = {"alpha": st["alpha"], "beta": st["beta"], "nu": st["nu"], "delta": st["delta"],
                  "phi": st["phi"], "gamma": st["gamma"], "pi": st["pi"]}
        code, _ = noalloc_bic(s, params)
        assert code == 1
        assert "use busseq_bic" in bq._load().busseq_last_error().decode()
        assert np.isfinite(s.bic(params))
