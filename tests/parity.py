"""Helpers of the GPU parity tests: the gfx950 sampler's state against the CPU
oracle's.  Both run the same Philox streams, numerics and reduction grouping,
so integer draws (X, Z, W, L) must match bit for bit and the continuous
parameters must be identical (rtol 1e-12 is written; equality is expected)."""
import os

import numpy as np

INT_KEYS = ("X", "Z", "W", "L")
FLOAT_KEYS = ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi", "ptau")


def oracle_threads():
    """OpenMP threads of the oracle: the host's cores, at most 16 (the CPU share of one GPU on a
    multi-GPU host)."""
    return min(16, os.cpu_count() or 1)


def compare(gs, os_, where):
    for k in INT_KEYS:
        a, b = gs[k], os_[k]
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{where}: {k} differs at {bad[:5].tolist()} ({len(bad)} entries) "
                                 f"gpu={a[tuple(bad[0])]} oracle={b[tuple(bad[0])]}")
    for k in FLOAT_KEYS:
        np.testing.assert_allclose(gs[k], os_[k], rtol=1e-12, atol=1e-300, err_msg=f"{where}: {k}")
