"""Benchmark of the BUSseq Gibbs sweep on one MI355X.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--config c2] [--precision 32]

One "step" is one full Gibbs sweep (W, pi, X/Z, alpha, gamma, L, p/tau0,
beta, nu, delta, phi) of the plain form: one chain on the synthetic data set
of the chosen configuration, counts and state resident in HBM.  Prints ONE JSON
line on stdout; diagnostics go to stderr.

Configurations (synthetic counts from the BUSseq generative model):
  c1  B=4, n_b=300/300/200/200 (N=1000), G=3000, K=5
  c2  B=4, n_b=8000/6000/4000/2000 (N=20000), G=3000, K=7   [default]
  c3  B=8, n_b=40k/30k/25k/20k/15k/10k/6k/4k (N=150000), G=5000, K=12, gamma0~U(-1,0.5)

The roofline leg times every kernel of the sweep with HIP events on the
sampler's own stream and reports the dominant one against HBM bandwidth.  The
cpu_baseline leg times the CPU oracle (OpenMP) on a bounded number of sweeps of
the same workload on this host.  Further legs (each skippable): c2 with FP64
per-element terms, c1 on the GPU beside the CPU oracle, the c4 K-sweep batch
and the c5 heterogeneous batch against their plain runs in sequence.

    python bench.py --batch-only c4|c5 --tile-rows TR [--steps K] [--warmup W]

times one batched sampler of c4 or c5 at a fixed tile size and prints its
kernel times instead of the benchmark line: the run the profiler passes over
the batched kernels use (one sampler per process, no autotuning).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from busseq_gpu_amd.workloads import CONFIGS, C4_CHAINS, C4_KS, c5_datasets  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Every timed sweep is a sweep of the recorded phase of a run (iterations past
# burn-in and past n_unchanged): the p/tau0 update and the posterior
# accumulation run on every sweep, as in the second half of a 4000-iteration run.
PHASE = dict(n_iter=4000, n_burnin=0, n_unchanged=0)
METRIC = "Gibbs sweeps/sec and cell×gene updates/s at N=20k,G=3k,K=7; batched vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, N, G, n_zero):
    """Algorithmic HBM bytes of one launch.  The streaming passes read the int32
    state matrix S once (4 B per cell x gene).  k_xz touches only the Y = 0
    entries: their uint16 gene index and their compact state word (6 B read),
    and writes the state to S and to the compact copy when it changes (at most
    8 B)."""
    if name == "k_xz":
        return 14.0 * n_zero
    if name in ("k_w", "k_alpha", "k_beta_nu", "k_delta", "k_phi"):
        return 4.0 * N * G
    return None


def pmc_profile(kernel, cfg_name, key):
    """A per-launch figure of the committed rocprofv3 PMC summary (HBM bytes per
    launch from FETCH_SIZE/WRITE_SIZE, or the VALU share of peak), if present."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{cfg_name}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("kernels", {}).get(kernel, {}).get(key)
    except (OSError, ValueError):
        return None


def cpu_baseline(Y, cfg, precision, budget_s):
    from oracle import pyoracle as po
    threads = min(16, os.cpu_count() or 1)
    o = po.Oracle(Y, cfg["n_b"], cfg["K"], seed=13579, precision=precision, threads=threads, **PHASE)
    t0 = time.perf_counter()
    o.sweeps(1)
    one = time.perf_counter() - t0
    n = max(1, min(2000, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.sweeps(n)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": n / dt, "unit": "Gibbs sweeps/s", "cores": threads, "kind": "port",
            "sample": f"{n} recorded-phase sweeps of the same workload after 1 warm-up sweep, CPU oracle "
                      f"(oracle/oracle.c, OpenMP, {threads} threads, precision {precision})"}


def plain_rate(bq, Y, cfg, precision, sweeps, warmup):
    """Plain form, one chain, tile size autotuned: (sweeps/s, tile rows, kernel ms)."""
    tr, _ = bq.autotune(Y, cfg["n_b"], cfg["K"], seed=13579, precision=precision, n_burnin=0, n_unchanged=0)
    with bq.Sampler(Y, cfg["n_b"], cfg["K"], seed=13579, precision=precision, tile_rows=tr, **PHASE) as s:
        s.sweeps(warmup)
        ms = s.time_sweeps(sweeps)
        kt = s.kernel_times(3)
    return sweeps / (ms / 1e3), tr, kt


def c1_leg(bq, precision, cpu_budget, no_cpu, sweeps=400, warmup=20):
    """c1 (B=4, N=1000, G=3000, K=5): the plain form on the GPU beside the CPU
    oracle on this host's cores, for the plain form's >= 30x target."""
    cfg = CONFIGS["c1"]
    Y = bq.simulate(cfg["n_b"], cfg["G"], cfg["K"], cfg["data_seed"], gamma0_range=cfg["g0"])
    gpu, tr, _ = plain_rate(bq, Y, cfg, precision, sweeps, warmup)
    out = {"config": "c1: N=1000 (300/300/200/200), G=3000, K=5, B=4", "sweeps": sweeps, "tile_rows": tr,
           "gpu_sweeps_per_s": round(gpu, 1), "cpu": None, "gpu_vs_cpu": None, "target": ">= 30x"}
    if not no_cpu:
        cpu = cpu_baseline(Y, cfg, precision, cpu_budget)
        out["cpu"] = cpu
        out["gpu_vs_cpu"] = round(gpu / cpu["value"], 1)
    return out


def fp64_leg(bq, Y, cfg, sweeps=100, warmup=5):
    """c2 with every per-element term in double precision (the reference's arithmetic)."""
    sps, tr, kt = plain_rate(bq, Y, cfg, 64, sweeps, warmup)
    return {"config": "c2, --precision 64 (per-element NB terms and lgamma in FP64)", "sweeps": sweeps,
            "tile_rows": tr, "sweeps_per_s": round(sps, 1),
            "cell_gene_updates_per_s": round(sps * sum(cfg["n_b"]) * cfg["G"], 1),
            "kernel_ms": {k: round(v, 5) for k, v in kt.items()}}


def c4_problems(bq):
    """c4: K = 3..10 with 4 chains each on the c1 data, as one list of problems."""
    cfg = CONFIGS["c1"]
    Y = bq.simulate(cfg["n_b"], cfg["G"], cfg["K"], cfg["data_seed"], gamma0_range=cfg["g0"])
    return [dict(Y=Y, n_b=cfg["n_b"], K=K, seed=13579, chains=C4_CHAINS, **PHASE) for K in C4_KS]


def c5_problems(bq):
    """c5: the 16 heterogeneous data sets, one chain each."""
    return [dict(Y=bq.simulate(d["n_b"], d["G"], d["K"], d["seed"]), n_b=d["n_b"], K=d["K"], seed=13579,
                 **PHASE) for d in c5_datasets()]


def batch_only(bq, which, precision, tile_rows, sweeps, warmup, kernel_sweeps):
    """One batched sampler of c4 or c5 at a fixed tile size, timed like the plain
    run (graph replays, then in-order launches with events between the kernels).
    A process that runs this creates one sampler and builds its graphs once,
    the shape of run the profiler passes over the batched kernels use."""
    probs = c4_problems(bq) if which == "c4" else c5_problems(bq)
    calib_ms = bq.copy_probe(1 << 30)
    log(f"[bench] copy probe 1 GiB: {calib_ms:.3f} ms = {2 * (1 << 30) / calib_ms / 1e6:.0f} GB/s (read+write)")
    cells = sum(int(p["Y"].size) * p.get("chains", 1) for p in probs)
    n_chains = sum(p.get("chains", 1) for p in probs)
    sb = bq.Sampler.batch(probs, precision=precision, tile_rows=tile_rows)
    sb.sweeps(warmup)
    ms = sb.time_sweeps(sweeps)
    kt = sb.kernel_times(kernel_sweeps)
    sb.close()
    rate = n_chains * sweeps / (ms / 1e3)
    log(f"[bench] batch {which}: {n_chains} chains, {sweeps} sweeps in {ms:.2f} ms -> {rate:.1f} chain-sweeps/s")
    for k, v in kt.items():
        log(f"[bench]   {k:14s} {v * 1e3:9.1f} us")
    return {"mode": "batch-only", "workload": which, "chains": n_chains, "tile_rows": tile_rows,
            "precision": precision, "sweeps": sweeps, "chain_sweeps_per_s": round(rate, 1),
            "cell_gene_updates_per_s": round(cells * sweeps / (ms / 1e3), 1),
            "kernel_ms": {k: round(v, 5) for k, v in kt.items()}}


def batched_vs_plain(bq, precision, sweeps=60, warmup=5):
    """c4: K-sweep K=3..10 with 4 chains each (32 chains) on the c1 data in one
    sampler (every chain in the same kernel launches) against the same
    32 chains run one after another in the plain form."""
    probs = c4_problems(bq)
    Y, n_b, G = probs[0]["Y"], probs[0]["n_b"], probs[0]["Y"].shape[1]
    Ks = list(C4_KS)
    chains = C4_CHAINS
    tr_b, _ = bq.autotune_batch(probs, precision=precision)
    sb = bq.Sampler.batch(probs, precision=precision, tile_rows=tr_b)
    sb.sweeps(warmup)
    ms_b = sb.time_sweeps(sweeps)
    sb.close()
    total_plain_ms = 0.0
    tr_p = {}
    for K in Ks:
        tr_p[K], _ = bq.autotune(Y, n_b, K, seed=13579, precision=precision, n_burnin=0, n_unchanged=0)
        for c in range(chains):
            with bq.Sampler(Y, n_b, K, seed=13579, chain0=c, precision=precision, tile_rows=tr_p[K],
                            **PHASE) as sp:
                sp.sweeps(warmup)
                total_plain_ms += sp.time_sweeps(sweeps)
    n_chains = len(Ks) * chains
    cs_b = n_chains * sweeps / (ms_b / 1e3)
    cs_p = n_chains * sweeps / (total_plain_ms / 1e3)
    N = sum(n_b)
    return {"config": "c4: K=3..10 x 4 chains on c1 data (N=1000, G=3000, B=4)", "chains": n_chains,
            "tile_rows_batched": tr_b, "tile_rows_plain": tr_p,
            "sweeps": sweeps, "batched_chain_sweeps_per_s": round(cs_b, 1),
            "plain_sequential_chain_sweeps_per_s": round(cs_p, 1), "ratio": round(cs_b / cs_p, 2),
            "batched_cell_gene_updates_per_s": round(cs_b * N * G, 1)}


def heterogeneous_batch(bq, precision, sweeps=20, warmup=3):
    """c5: the 16 data sets in one batched sampler (every data set in the same launches)
    against the same 16 chains run one after another in the plain form."""
    sets = c5_datasets()
    probs = c5_problems(bq)
    cells = sum(sum(d["n_b"]) * d["G"] for d in sets)
    tr_b, _ = bq.autotune_batch(probs, precision=precision, sweeps=5)
    sb = bq.Sampler.batch(probs, precision=precision, tile_rows=tr_b)
    sb.sweeps(warmup)
    ms_b = sb.time_sweeps(sweeps)
    sb.close()
    ms_p = 0.0
    for p in probs:
        tr, _ = bq.autotune(p["Y"], p["n_b"], p["K"], seed=13579, precision=precision, sweeps=5,
                            n_burnin=0, n_unchanged=0)
        with bq.Sampler(p["Y"], p["n_b"], p["K"], seed=13579, precision=precision, tile_rows=tr,
                        **PHASE) as sp:
            sp.sweeps(warmup)
            ms_p += sp.time_sweeps(sweeps)
    return {"config": "c5: 16 data sets, N 500-30000 log-uniform, G 1000-5000, B 2-6, K 3-10",
            "cells_x_genes": cells, "total_cells": sum(sum(d["n_b"]) for d in sets), "sweeps": sweeps,
            "tile_rows_batched": tr_b,
            "batched_chain_sweeps_per_s": round(len(sets) * sweeps / (ms_b / 1e3), 1),
            "plain_sequential_chain_sweeps_per_s": round(len(sets) * sweeps / (ms_p / 1e3), 1),
            "ratio": round(ms_p / ms_b, 2),
            "batched_cell_gene_updates_per_s": round(cells * sweeps / (ms_b / 1e3), 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--kernel-sweeps", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the c1 and FP64 legs")
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="cells per tile of the plain run (0: the autotuner's choice); "
                         "fixed, it keeps separate profiler passes on the same tiles")
    ap.add_argument("--batch-only", choices=("c4", "c5"), default=None,
                    help="time one batched sampler of c4 or c5 at --tile-rows and print its kernel "
                         "times instead of the benchmark line (for profiler passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(chains, K-sweeps and data sets shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)

    import busseq_gpu_amd as bq

    if args.batch_only:
        if args.tile_rows <= 0:
            print("bench.py: --batch-only needs --tile-rows", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(batch_only(bq, args.batch_only, args.precision, args.tile_rows, args.steps,
                                    args.warmup, args.kernel_sweeps)), flush=True)
        return

    cfg = CONFIGS[args.config]
    n_b, G, K = cfg["n_b"], cfg["G"], cfg["K"]
    N = sum(n_b)
    t0 = time.perf_counter()
    Y = bq.simulate(n_b, G, K, cfg["data_seed"], gamma0_range=cfg["g0"])
    n_zero = int((Y == 0).sum())
    log(f"[bench] data {args.config}: N={N} G={G} K={K} zeros={n_zero / Y.size:.3f} ({time.perf_counter() - t0:.1f}s)")

    calib_ms = bq.copy_probe(1 << 30)
    log(f"[bench] copy probe 1 GiB: {calib_ms:.3f} ms = {2 * (1 << 30) / calib_ms / 1e6:.0f} GB/s (read+write)")

    if args.tile_rows > 0:
        tr, how = args.tile_rows, "given"
    else:
        tr, _ = bq.autotune(Y, n_b, K, seed=13579, precision=args.precision, n_burnin=0, n_unchanged=0)
        how = "autotuned"
    t0 = time.perf_counter()
    s = bq.Sampler(Y, n_b, K, seed=13579, precision=args.precision, tile_rows=tr, **PHASE)
    log(f"[bench] init {time.perf_counter() - t0:.1f}s, tile rows {tr} ({how})")
    s.sweeps(args.warmup)
    w0 = time.perf_counter()
    ms = s.time_sweeps(args.steps)
    wall = time.perf_counter() - w0
    sweeps_per_s = args.steps / (ms / 1e3)
    log(f"[bench] {args.steps} sweeps: {ms:.2f} ms (events), wall {wall * 1e3:.2f} ms -> {sweeps_per_s:.1f} sweeps/s")

    kt = s.kernel_times(args.kernel_sweeps)
    for k, v in kt.items():
        log(f"[bench]   {k:14s} {v * 1e3:9.1f} us")
    dom = max(kt, key=kt.get)
    dom_bytes = kernel_bytes(dom, N, G, n_zero)
    roof = None
    if dom_bytes is not None:
        ach = dom_bytes / (kt[dom] * 1e-3) / 1e9
        traffic = pmc_profile(dom, args.config, "bytes_per_launch")
        valu = pmc_profile(dom, args.config, "valu_share")
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                "traffic": traffic, "algorithmic_bytes": dom_bytes, "avg_ms": round(kt[dom], 5),
                "valu_share_profiled": round(valu, 3) if valu is not None else None}
    sweep_bytes = 16.0 * N * G
    s.close()

    batched = None
    if not args.no_batched:
        batched = batched_vs_plain(bq, args.precision)
        log(f"[bench] batched c4: {batched['batched_chain_sweeps_per_s']} chain-sweeps/s vs plain sequential "
            f"{batched['plain_sequential_chain_sweeps_per_s']} -> {batched['ratio']}x")

    hetero = None
    if not args.no_batched and not args.no_c5:
        hetero = heterogeneous_batch(bq, args.precision)
        log(f"[bench] batched c5: {hetero['batched_chain_sweeps_per_s']} chain-sweeps/s vs plain sequential "
            f"{hetero['plain_sequential_chain_sweeps_per_s']} -> {hetero['ratio']}x "
            f"({hetero['batched_cell_gene_updates_per_s']:.3g} cell x gene updates/s)")

    fp64 = None
    if not args.no_legs and args.precision == 32 and args.config == "c2":
        fp64 = fp64_leg(bq, Y, cfg)
        log(f"[bench] c2 FP64 terms: {fp64['sweeps_per_s']} sweeps/s (tile rows {fp64['tile_rows']})")

    c1 = None
    if not args.no_legs and args.config != "c1":
        c1 = c1_leg(bq, args.precision, args.cpu_budget / 2, args.no_cpu)
        log(f"[bench] c1: GPU {c1['gpu_sweeps_per_s']} sweeps/s, CPU oracle "
            f"{c1['cpu']['value'] if c1['cpu'] else float('nan'):.2f} -> {c1['gpu_vs_cpu']}x")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(Y, cfg, args.precision, args.cpu_budget)
        log(f"[bench] cpu oracle: {cpu['value']:.4f} sweeps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(sweeps_per_s, 3),
        "unit": "Gibbs sweeps/s",
        "cell_gene_updates_per_s": round(sweeps_per_s * N * G, 1),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic (BUSseq generative model, seeded)",
        "config": {"workload": f"plain BUSseq Gibbs sweep, one chain, {args.config}: N={N} "
                               f"({'/'.join(map(str, n_b))}), G={G}, K={K}, B={len(n_b)}",
                   "chains": 1, "precision": args.precision, "tile_rows": tr, "parallelism": "single GPU",
                   "phase": "recorded phase: p/tau0 update and posterior accumulation on every sweep"},
        "roofline": roof,
        "sweep_roofline": {"bytes_per_sweep": sweep_bytes,
                           "achieved": round(sweep_bytes * sweeps_per_s / 1e9, 2),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(sweep_bytes * sweeps_per_s / 1e9 / HBM_PEAK_GBS, 5)},
        "kernel_ms": {k: round(v, 5) for k, v in kt.items()},
        "batched_vs_plain": batched,
        "heterogeneous_batch": hetero,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(sweeps_per_s / cpu["value"], 1) if cpu else None,
        "c1_plain_vs_cpu": c1,
        "c2_fp64": fp64,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
