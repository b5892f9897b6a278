#!/usr/bin/env python3
"""Benchmark: aggregate Metropolis trial moves per second on one MI355X.

Workload (one "step"): 1024 semi-grand-canonical chains on a binary FCC
10x10x10 primitive supercell (1000 sites; pairs to the 4th shell, triplets to
the 2nd), 64 temperatures log-spaced over 300-1500 K x 8 chemical potentials
in [-0.2, 0.2] eV x 2 seeds, each chain advanced by 500 sweeps (5e5 flips) in
one persistent launch.  Beside it: the plain path (the same kernel with one
chain per launch, canonical swaps at 800 K, 2e5 moves), the roofline of the
sweep kernel from HIP events, and the CPU oracle timed on the host cores.

Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aggregate MC trial moves/sec at 1 MI355X (batched vs plain); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes of one SGC launch of the default workload (1024 chains x 500
# sweeps): FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
# counter reads for a copy kernel of known size (x2.0 and x1.0).  Measured
# with rocprofv3 --pmc; summary in profiles/sweep_kernel_sgc_binary_fcc10_1024chains.txt
TRAFFIC_DEFAULT_LAUNCH = 2491371


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def binary_model():
    from cemc_amd.lattice import FCCSettings, seeded_ecis
    from cemc_amd.model import CEModel
    s = FCCSettings(10, ["A", "B"], {2: 4, 3: 2})
    eci = seeded_ecis(s, 2024, 0.02)
    return s, eci, CEModel(s, eci)


def batched_chains(settings, n_chains=1024):
    from cemc_amd.lattice import random_occupation
    temps = np.geomspace(300.0, 1500.0, 64)
    mus = np.linspace(-0.2, 0.2, 8)
    chains = []
    for c in range(n_chains):
        t = temps[c % 64]
        mu = mus[(c // 64) % 8]
        seed_rep = c // 512
        chains.append(dict(kind=1, temperature=float(t), chem_pot={"c1_0": float(mu)},
                           occupation=random_occupation(settings.n_sites, 2, 1000 + c),
                           seed=77 + seed_rep, uid=c))
    return chains


def cpu_baseline(settings, eci, chains, target_s=12.0):
    """Oracle (C restatement) chains on the host cores, one chain per thread."""
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater, run_many
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    names = settings.cf_names()
    model = OracleModel(settings, names)
    ups, uids, temps = [], [], []
    for c in range(threads):
        ch = chains[c]
        e = dict(eci)
        e["c1_0"] -= ch["chem_pot"]["c1_0"]
        cf0 = brute_force_cf(settings, ch["occupation"], names)
        ups.append(OracleUpdater(model, ch["occupation"], cf0, e))
        uids.append(ch["uid"])
        temps.append(ch["temperature"])
    # calibrate on a short run, then time a bounded sample
    t0 = time.perf_counter()
    run_many(ups, 1, 77, uids, 2000, temps, threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    per_thread = int(max(2000, min(2000 * target_s / dt, 5e7)))
    t0 = time.perf_counter()
    run_many(ups, 1, 77, uids, per_thread, temps, threads, move_begin=2000)
    dt = time.perf_counter() - t0
    return {"value": threads * per_thread / dt, "unit": "moves/s", "cores": threads, "kind": "port",
            "sample": "%d independent SGC chains of the same workload (one per thread), %d flips each, %.1f s"
                      % (threads, per_thread, dt)}


def copy_probe(mib):
    """One copy kernel that reads and writes `mib` MiB each: the known traffic
    against which the HBM byte counters are scaled."""
    from cemc_amd import probe
    ms, gbs = probe.hbm_copy(mib << 20, reps=1)
    log("copy probe: %d MiB read + %d MiB written in %.3f ms (%.0f GB/s)" % (mib, mib, ms, gbs))


def measured_peaks():
    """HBM copy and LDS read rates measured on this GPU by the probe kernels."""
    from cemc_amd import probe
    _, hbm = probe.hbm_copy(1 << 30, reps=10)
    _, lds = probe.lds_read()
    log("measured peaks: HBM copy %.0f GB/s, LDS read %.0f GB/s" % (hbm, lds))
    return hbm, lds


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=500)
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this size first (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: each chain lives on one device and this build runs "
              "on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cemc_amd import _capi
    from cemc_amd.batch import ChainBatch
    from cemc_amd.lattice import random_occupation
    _capi.require_gpu()

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    hbm_meas, lds_meas = measured_peaks()
    settings, eci, model = binary_model()
    log("model: periodic=%s sub-clusters=%d gathers/flip=%d" % (model.periodic, model.n_subclusters,
                                                                model.gathers_per_flip()))
    chains = batched_chains(settings, args.chains)
    batch = ChainBatch([model] * len(chains), chains)
    moves_per_chain = args.sweeps * settings.n_sites
    for _ in range(args.warmup):
        batch.run(moves_per_chain)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(batch.run_timed(moves_per_chain))
    wall = time.perf_counter() - t0
    total_moves = args.steps * moves_per_chain * len(chains)
    value = total_moves / wall
    ms_step = 1e3 * wall / args.steps
    avg_kernel_ms = float(np.mean(kernel_ms))
    st = batch.stats()
    log("batched: %.3e moves/s, kernel %.2f ms/step, acceptance %.3f" % (
        value, avg_kernel_ms, st["accepted"].sum() / max(st["moves"].sum(), 1)))

    # plain path: one chain per launch (canonical swaps at 800 K, 2e5 moves)
    occ = random_occupation(settings.n_sites, 2, 1)
    plain = ChainBatch([model], [dict(kind=0, temperature=800.0, occupation=occ, seed=42, uid=0)])
    plain.run(2000)
    t0 = time.perf_counter()
    plain_ms = plain.run_timed(200000)
    plain_wall = time.perf_counter() - t0
    plain_value = 200000 / plain_wall
    log("plain: %.3e moves/s (kernel %.2f ms)" % (plain_value, plain_ms))

    bytes_per_flip = 5.0 * model.gathers_per_flip()
    launch_bytes = bytes_per_flip * moves_per_chain * len(chains)
    achieved = launch_bytes / (avg_kernel_ms * 1e-3) / 1e9
    default_shape = len(chains) == 1024 and args.sweeps == 500
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_DEFAULT_LAUNCH if default_shape else None,
                "lds_peak_measured": lds_meas, "lds_frac": achieved / lds_meas,
                "hbm_copy_measured": hbm_meas,
                "bytes_per_move": bytes_per_flip, "kernel_ms": avg_kernel_ms}
    out = {
        "metric": METRIC, "value": value, "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched SGC, binary FCC 10x10x10 (1000 sites), pairs<=4th + triplets<=2nd shell, "
                               "%d chains (64 T in 300-1500 K x 8 mu in [-0.2,0.2] eV x 2 seeds), %d sweeps "
                               "per step, one persistent launch" % (len(chains), args.sweeps),
                   "chains": len(chains), "sweeps_per_step": args.sweeps, "sites": settings.n_sites},
        "plain_moves_per_s": plain_value, "batched_over_plain": value / plain_value,
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(settings, eci, chains)
        log("cpu baseline: %.3e moves/s on %d threads" % (out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
