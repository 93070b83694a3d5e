"""The benchmark workloads: synthetic data sets from the BUSseq generative model
at the shapes the sampler is measured on, shared by bench.py and the tests.

    c1  B=4, n_b=300/300/200/200 (N=1000), G=3000, K=5
    c2  B=4, n_b=8000/6000/4000/2000 (N=20000), G=3000, K=7
    c3  B=8, n_b=40k/30k/25k/20k/15k/10k/6k/4k (N=150000), G=5000, K=12, gamma0 ~ U(-1, 0.5)
    c4  the c1 data with K = 3..10, 4 chains each (32 chains)
    c5  16 data sets: N log-uniform in [500, 30000], G uniform in [1000, 5000],
        B in {2..6}, K in {3..10}, batch sizes Dir(2) N with at least K + 1
        cells per batch

Data seed 1000 + configuration number (c1..c3), 5000 + index (c5); the
sampler's seed is 13579 throughout (README.md:21 of the reference).
"""
import numpy as np

CONFIGS = {
    "c1": dict(n_b=[300, 300, 200, 200], G=3000, K=5, data_seed=1001, g0=(-2.0, 0.0)),
    "c2": dict(n_b=[8000, 6000, 4000, 2000], G=3000, K=7, data_seed=1002, g0=(-2.0, 0.0)),
    "c3": dict(n_b=[40000, 30000, 25000, 20000, 15000, 10000, 6000, 4000], G=5000, K=12,
               data_seed=1003, g0=(-1.0, 0.5)),
}
SEED = 13579
C4_KS = tuple(range(3, 11))
C4_CHAINS = 4


def counts(bq, name):
    """Y of configuration c1, c2 or c3."""
    cfg = CONFIGS[name]
    return bq.simulate(cfg["n_b"], cfg["G"], cfg["K"], cfg["data_seed"], gamma0_range=cfg["g0"])


def c4_problems(bq, Y=None, **phase):
    """c4: the K-sweep K = 3..10 with 4 chains each on the c1 data, as the
    problem list of one batched sampler (chain j of problem K has key word j)."""
    cfg = CONFIGS["c1"]
    if Y is None:
        Y = counts(bq, "c1")
    return [dict(Y=Y, n_b=cfg["n_b"], K=K, seed=SEED, chains=C4_CHAINS, **phase) for K in C4_KS]


def c5_datasets(n=16):
    """c5: the shapes of the 16 heterogeneous data sets (data seed 5000 + index)."""
    out = []
    for j in range(n):
        rng = np.random.default_rng(5000 + j)
        N = int(round(np.exp(rng.uniform(np.log(500), np.log(30000)))))
        G = int(rng.integers(1000, 5001))
        B = int(rng.integers(2, 7))
        K = int(rng.integers(3, 11))
        N = max(N, B * (K + 1))
        extra = rng.multinomial(N - B * (K + 1), rng.dirichlet([2.0] * B))
        out.append(dict(n_b=[int(K + 1 + e) for e in extra], G=G, K=K, seed=5000 + j))
    return out


def c5_problems(bq, **phase):
    """c5 as the problem list of one batched sampler, one chain per data set."""
    probs = []
    for d in c5_datasets():
        Y = bq.simulate(d["n_b"], d["G"], d["K"], d["seed"])
        probs.append(dict(Y=Y, n_b=d["n_b"], K=d["K"], seed=SEED, **phase))
    return probs
