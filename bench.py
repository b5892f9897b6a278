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
# sweeps, 1 wave per chain, rounds of 8 moves): FETCH_SIZE and WRITE_SIZE per
# dispatch, each scaled by what the counter reads for a copy kernel of known
# size (x2.0 and x1.0).  Measured with rocprofv3 --pmc; summary in
# profiles/sweep_kernel_sgc_binary_fcc10_1024chains.txt
TRAFFIC_DEFAULT_LAUNCH = 3686400
TRAFFIC_SHAPE = (1, 8, 1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# What bounds the sweep kernel: neither bandwidth (the algorithmic bytes are
# a few % of the measured LDS read rate and the lattices never leave LDS
# during a launch) but each chain's dependent move sequence (evaluate ->
# reduce -> decide -> commit), which the profiles' SQ counters show as a low
# VALU issue fraction and waves that wait on their LDS and HBM table reads.
LIMITER = ("latency of each chain's dependent move sequence: per-chain issue and LDS/L2 latency, not HBM or "
           "LDS bandwidth (SQ counters in profiles/)")


def roofline_of(algorithmic_bytes, kernel_ms, lds_peak):
    """Roofline object of one launch: algorithmic bytes / launch time against
    HBM's published peak, and against the LDS read rate measured here."""
    achieved = algorithmic_bytes / (kernel_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "limiter": LIMITER}
    if lds_peak:
        out["lds_peak_measured"] = lds_peak
        out["lds_frac"] = achieved / lds_peak
    return out


def oracle_sample(models, chains, occs, cfs, kind, seconds, what):
    """CPU baseline: the oracle (C restatement) on up to 16 host threads, one
    chain each, from the given lattices (their correlation functions seed
    the oracle), for about `seconds`."""
    from oracle.oracle import OracleModel, OracleUpdater, run_many
    threads = max(1, min(len(models), int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 16))
    models, chains, occs, cfs = models[:threads], chains[:threads], occs[:threads], cfs[:threads]
    omodels = {}
    ups, uids, temps = [], [], []
    for m, ch, occ, cf in zip(models, chains, occs, cfs):
        if id(m) not in omodels:
            omodels[id(m)] = OracleModel(m.settings, m.names)
        eci = dict(zip(m.names, np.asarray(m.eci).tolist()))
        for name, mu in (ch.get("chem_pot") or {}).items():
            eci[name] -= mu
        ups.append(OracleUpdater(omodels[id(m)], occ, dict(zip(m.names, np.asarray(cf).tolist())), eci))
        uids.append(ch["uid"])
        temps.append(ch["temperature"])
    seed = chains[0]["seed"]
    n = 64
    t0 = time.perf_counter()
    run_many(ups, kind, seed, uids, n, temps, threads, move_begin=1 << 40)
    dt = max(time.perf_counter() - t0, 1e-4)
    n = int(max(64, min(n * seconds / dt, 5e7)))
    t0 = time.perf_counter()
    run_many(ups, kind, seed, uids, n, temps, threads, move_begin=(1 << 40) + 64)
    dt = time.perf_counter() - t0
    return {"value": threads * n / dt, "unit": "moves/s", "cores": threads, "kind": "port",
            "sample": "%s, %d moves each, %.1f s" % (what, n, dt)}


def reference_sample(kind, seconds, procs):
    """The reference's own CEUpdater (oracle/_ref, built from the reference's
    sources) driven one move at a time by the fed-stream sampler loop, in
    `procs` parallel processes; None when its interpreter or build is absent."""
    import subprocess
    py39 = "/opt/conda/bin/python3.9"
    so_dir = os.path.join(ROOT, "oracle", "_ref")
    if not (os.path.exists(py39) and os.path.isdir(so_dir) and any(f.startswith("cemc_ref") for f in os.listdir(so_dir))):
        return None
    try:
        res = subprocess.run([py39, os.path.join(ROOT, "oracle", "ref_rate.py"), kind, str(seconds), str(procs)],
                             capture_output=True, text=True, timeout=seconds * 6 + 120)
    except subprocess.TimeoutExpired:
        return None
    if res.returncode != 0:
        log("reference sample failed: " + res.stderr[-400:])
        return None
    return json.loads(res.stdout.strip().splitlines()[-1])


def dropin_rate(settings, eci, seconds=3.0):
    """The drop-in path on the GPU: cemc_amd.updater.PyCEUpdater driven one
    move at a time by the same fed-stream loop the reference's samplers use
    (one calculate, then clear_history or undo_changes, per move), on the
    plain configuration (canonical swaps at 800 K); with the resident
    updater kernel (the default) and with one kernel launch per call."""
    from cemc_amd.dropin import run_updater_chain
    from cemc_amd.lattice import SimpleAtoms, random_occupation
    from cemc_amd.updater import PyCEUpdater

    def rate(resident):
        occ = random_occupation(settings.n_sites, 2, 1)
        atoms = SimpleAtoms([settings.species[v] for v in occ])
        upd = PyCEUpdater(atoms, settings, None, dict(eci))
        upd.set_resident(resident)
        n = 200
        t0 = time.perf_counter()
        run_updater_chain(upd, atoms, settings.species, 0, 42, 0, n, 800.0)
        dt = time.perf_counter() - t0
        n = int(max(200, n * seconds / max(dt, 1e-3)))
        t0 = time.perf_counter()
        _, _, _, _, acc = run_updater_chain(upd, atoms, settings.species, 0, 42, 1, n, 800.0)
        dt = time.perf_counter() - t0
        upd.close()
        return n / dt, n, acc, dt

    v, n, acc, dt = rate(True)
    v1, n1, _, _ = rate(False)
    return {"moves_per_s": v, "moves": n, "accepted": acc, "seconds": dt,
            "launch_per_call_moves_per_s": v1, "launch_per_call_moves": n1,
            "workload": "PyCEUpdater on binary FCC 10^3, canonical swaps at 800 K, one calculate + "
                        "clear_history/undo_changes per move from a Python loop (the reference's per-move protocol); "
                        "calls as commands to the resident updater kernel (launch_per_call_*: one launch per call)"}


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


def ternary_uneven(n_chains=256, sweeps=200, seed=3, lds_peak=None, cpu=True):
    """256 canonical chains, ternary FCC with 2 polynomial basis functions,
    pairs to the 4th shell, triplets to the 3rd, quadruplets to the 2nd,
    ECIs N(0, 0.01 eV)/d^2; cells L^3 with L uniform in 8..16 (512-4096
    sites), compositions Dirichlet(1,1,1), T uniform in 300-1500 K; every
    chain makes `sweeps` sweeps in one launch."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    rng = np.random.Generator(np.random.Philox(seed))
    sizes = rng.integers(8, 17, n_chains)
    temps = rng.uniform(300.0, 1500.0, n_chains)
    fracs = rng.dirichlet([1.0, 1.0, 1.0], n_chains)
    models = {}
    chains, chain_models = [], []
    for c in range(n_chains):
        L = int(sizes[c])
        if L not in models:
            st = FCCSettings(L, ["A", "B", "C"], {2: 4, 3: 3, 4: 2})
            models[L] = (st, CEModel(st, seeded_ecis(st, 2025, 0.01)))
        st, model = models[L]
        f = np.maximum(fracs[c], 0.02)
        f = f / f.sum()
        occ = random_occupation(st.n_sites, 3, 5000 + c, fractions=list(f))
        chains.append(dict(kind=CANONICAL, temperature=float(temps[c]), occupation=occ, seed=500, uid=c))
        chain_models.append(model)
    from cemc_amd import autotune
    tune_w, tune_rates = autotune.tune(chain_models, chains, candidates=(1, 2, 4, 8), sweeps=1.0, log=log)
    batch = ChainBatch(chain_models, chains)
    autotune.apply(batch, tune_w)
    moves = np.array([sweeps * m.n_sites for m in chain_models], dtype=np.int64)
    batch.run(moves // 20)  # warm-up: a tenth of a step
    t0 = time.perf_counter()
    kernel_ms = batch.run_timed(moves)
    wall = time.perf_counter() - t0
    shapes = sorted({batch.launch_shape(c) for c in range(n_chains)})
    total = int(moves.sum())
    m0 = models[int(sizes[0])][1]
    bytes_per_move = 10.0 * m0.gathers_per_flip()
    out = {"workload": "ternary FCC, 256 canonical chains, L^3 with L in 8..16 (512-4096 sites), pairs<=4th, "
                       "triplets<=3rd, quadruplets<=2nd shell, %d sweeps each, one launch" % sweeps,
           "value": total / wall, "unit": "moves/s", "moves": total, "kernel_ms": kernel_ms, "wall_s": wall,
           "bytes_per_move": bytes_per_move, "launch_shapes": [list(x) for x in shapes],
           "shape": autotune.shape_name(tune_w),
           "autotune_moves_per_s": {autotune.shape_name(k): v for k, v in tune_rates.items()},
           "acceptance": float(batch.stats()["accepted"].sum() / max(total + int((moves // 20).sum()), 1)),
           "roofline": roofline_of(bytes_per_move * total, kernel_ms, lds_peak)}
    if cpu:
        # 16 of the chains (one per host thread) from their current lattices
        pick = list(range(0, n_chains, max(1, n_chains // 16)))[:16]
        out["cpu_baseline"] = oracle_sample([chain_models[c] for c in pick], [chains[c] for c in pick],
                                            [batch.occupation(c) for c in pick], [batch.cf(c)[0] for c in pick],
                                            kind=CANONICAL, seconds=6.0,
                                            what="16 of the 256 chains (one per thread, sizes as drawn)")
    batch.close()
    return out


def tempering(n_rep=128, cycles=2000, L=32, seed=9, lds_peak=None, cpu=True):
    """Quaternary FCC 32^3 (32768 sites), 3 basis functions, pairs,
    triplets and quadruplets to the 3rd shell; 128 replicas on
    geomspace(200, 2000) K; one canonical sweep per replica then one
    on-device exchange step per cycle; cycles replayed from HIP graphs."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    from cemc_amd.tempering import ReplicaExchange
    st = FCCSettings(L, ["A", "B", "C", "D"], {2: 3, 3: 3, 4: 3})
    model = CEModel(st, seeded_ecis(st, 2026, 0.02))
    temps = np.geomspace(200.0, 2000.0, n_rep)
    chains = [dict(kind=CANONICAL, temperature=float(T), occupation=random_occupation(st.n_sites, 4, 7000 + r),
                   seed=seed, uid=r) for r, T in enumerate(temps)]
    from cemc_amd import autotune
    tune_w, tune_rates = autotune.tune([model] * n_rep, chains, candidates=(2, 4, 8), sweeps=0.25, repeats=1,
                                       log=log)
    batch = ChainBatch([model] * n_rep, chains)
    autotune.apply(batch, tune_w)
    pt = ReplicaExchange(batch, seed=seed, uid=1 << 40)
    pt.run(2, st.n_sites)  # warm-up and graph build
    # timed in pieces of at most 200 cycles (every cycle is the same sweep and
    # exchange), with a progress line after each piece
    t0 = time.perf_counter()
    kernel_ms, done = 0.0, 0
    while done < cycles:
        n = min(200, cycles - done)
        kernel_ms += pt.run(n, st.n_sites)
        done += n
        if cycles > 200:
            log("tempering: %d of %d cycles, %.1f s" % (done, cycles, time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    state = pt.state()
    total = cycles * n_rep * st.n_sites
    acc = state["n_accept"].sum() / max(state["n_try"].sum(), 1)
    bytes_per_move = 10.0 * model.gathers_per_flip()
    out = {"workload": "quaternary FCC 32^3 (32768 sites), <=4-body to the 3rd shell (%d CFs, %d sub-clusters per "
                       "site), %d replicas 200-2000 K, exchange every sweep, %d of the workload's 2000 sweeps%s"
                       % (model.n_cf, model.n_subclusters, n_rep, cycles,
                          "" if cycles >= 2000 else " (a rate per cycle: every cycle is the same sweep + exchange)"),
           "cycles": cycles, "value": total / wall, "unit": "moves/s", "moves": total, "kernel_ms": kernel_ms,
           "wall_s": wall, "bytes_per_move": bytes_per_move, "launch_shape": list(batch.launch_shape(0)),
           "shape": autotune.shape_name(tune_w),
           "autotune_moves_per_s": {autotune.shape_name(k): v for k, v in tune_rates.items()},
           "exchange_accept_rate": float(acc),
           "roofline": roofline_of(bytes_per_move * total, kernel_ms, lds_peak)}
    if cpu:
        pick = list(range(0, n_rep, n_rep // 16))[:16]
        out["cpu_baseline"] = oracle_sample([model] * len(pick), [chains[c] for c in pick],
                                            [batch.occupation(c) for c in pick], [batch.cf(c)[0] for c in pick],
                                            kind=CANONICAL, seconds=6.0,
                                            what="16 of the 128 replicas (one per thread), sweeps without exchange")
    pt.close()
    batch.close()
    return out


def wang_landau(model, settings, n_walkers=1024, n_checks=4, check_every=5000):
    """Wang-Landau walkers on the binary FCC 10^3 model: one canonical walker
    per wave, 100 bins over +-2 eV around each start, f = 1; then the same
    walkers on one shared histogram."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.lattice import random_occupation
    chains = [dict(kind=CANONICAL, temperature=1000.0, occupation=random_occupation(settings.n_sites, 2, 5000 + c),
                   seed=91, uid=c) for c in range(n_walkers)]
    b = ChainBatch([model] * n_walkers, chains)
    e = b.energies()
    for c in range(n_walkers):
        b.wl_set(c, 100, e[c] - 2.0, e[c] + 2.0, 1.0, 1e-8, 0.8, current_bin=50)
    b.wl_run(1, 1000)
    _, ms = b.wl_run(n_checks, check_every, timed=True)
    moves = n_walkers * n_checks * check_every
    w0 = b.wl_get(0)
    # the same walkers on one shared histogram spanning every start
    lo, hi = float(e.min()) - 2.0, float(e.max()) + 2.0
    for c in range(n_walkers):
        b.wl_set(c, 100, lo, hi, 1.0, 1e-8, 0.8)
    b.wl_run(1, 1000 * n_walkers, shared=True)
    _, ms_sh = b.wl_run(n_checks, check_every * n_walkers, timed=True, shared=True)
    b.close()
    return {"value": moves / (ms * 1e-3), "unit": "swaps/s", "kernel_ms": ms, "walkers": n_walkers,
            "moves": moves, "bins": 100, "acceptance_walker0": w0["n_accept"] / max(w0["n_moves"], 1),
            "shared_histogram": {"value": moves / (ms_sh * 1e-3), "unit": "swaps/s", "kernel_ms": ms_sh,
                                 "workload": "the same walkers and moves on one histogram of 100 bins over all "
                                             "their starts +-2 eV, one launch and one flatness test per check"}}


def serving(n_jobs=5000, rate=200.0):
    """Poisson stream of MC jobs (binary or ternary FCC, L^3 with L in 6..24,
    canonical or SGC, 100-5000 sweeps) through the continuous-batching job
    server; latency = completion - arrival on the loop's wall clock."""
    from cemc_amd.serving import JobServer, synthetic_stream
    cache = {}
    t0 = time.perf_counter()
    jobs = synthetic_stream(n_jobs=n_jobs, rate=rate, seed=4, model_cache=cache)
    setup = time.perf_counter() - t0
    server = JobServer(max_chains=4096, hbm_bytes=1 << 30)
    res = server.serve(jobs, realtime=True, progress=log)
    moves = np.array([j.n_moves for j in jobs], dtype=np.float64)
    span = max(j.arrival for j in jobs)
    res.update({"workload": "%d Poisson jobs at %.0f/s: binary or ternary FCC L^3, L in 6..24, canonical or SGC, "
                            "T in 300-1500 K, 100-5000 sweeps; continuous batching on one GPU" % (n_jobs, rate),
                "setup_s": setup, "models": len(cache), "rejected": len(server.rejected),
                # the load the stream offers and the arrival rate the measured throughput would sustain
                "offered_moves_per_s": float(moves.sum() / span),
                "sustainable_jobs_per_s": res["throughput_moves_per_s"] / float(moves.mean())})
    return res


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
    # reference-shaped: one chain, the flip's CF updates split over threads
    # (the reference's set_num_threads, ce_updater.cpp:352); config-1 swaps
    from cemc_amd.lattice import random_occupation
    occ1 = random_occupation(settings.n_sites, 2, 42)
    shaped = {}
    for t in (1, 2, 4, 8, 16):
        if t > threads:
            break
        up = OracleUpdater(model, occ1, brute_force_cf(settings, occ1, names), dict(eci))
        up.set_num_threads(t)
        n1 = 2000
        t1 = time.perf_counter()
        up.run(0, 42, 0, n1, 800.0)
        d1 = max(time.perf_counter() - t1, 1e-3)
        n1 = int(max(2000, min(n1 * 1.5 / d1, 5e6)))
        t1 = time.perf_counter()
        up.run(0, 42, 0, n1, 800.0, move_begin=2000)
        shaped[str(t)] = n1 / (time.perf_counter() - t1)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    port = {"value": threads * per_thread / dt, "unit": "moves/s", "cores": threads, "kind": "port",
            "sample": "%d independent SGC chains of the same workload (one per thread), %d flips each, %.1f s"
                      % (threads, per_thread, dt),
            "reference_shaped": {"workload": "one canonical chain of the plain configuration (binary FCC 10^3, "
                                             "800 K), CF updates of each flip split over T threads; swaps/s by T",
                                 "moves_per_s": shaped}}
    ref = reference_sample("sgc", 8.0, threads)
    if ref is None:
        port.update({"cpu_model": cpu_model, "nproc": os.cpu_count()})
        return port
    return {"value": ref["moves_per_s"], "unit": "moves/s", "cores": ref["processes"], "kind": "reference",
            "sample": "the reference's CEUpdater (compiled from its sources) in %d processes, one SGC chain of the "
                      "workload each (its T and mu), driven one flip at a time from Python as its samplers do; "
                      "%d flips in %.1f s" % (ref["processes"], ref["moves"], ref["wall_s"]),
            "port": port, "cpu_model": cpu_model, "nproc": os.cpu_count()}


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
    ap.add_argument("--workloads", default="batched,dropin,ternary,tempering,serving,wang_landau",
                    help="comma list of batched, dropin, ternary, tempering, serving, wang_landau")
    ap.add_argument("--shape", default=None,
                    help="'fields' (per-site fields, one wave per chain) or WxK (waves per chain, moves per round, "
                         "tables summed per move) for the batched workload; skips the autotuner")
    ap.add_argument("--pt-cycles", type=int, default=200)
    ap.add_argument("--serving-jobs", type=int, default=500)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this size first (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: each chain lives on one device and the engine runs "
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
    from cemc_amd import autotune
    if args.shape:
        tune_w, tune_rates = autotune.parse_shape(args.shape), {}
    else:
        tune_w, tune_rates = autotune.tune([model] * len(chains), chains, candidates=(1, 2, 4), sweeps=20.0,
                                           repeats=1, log=log)
    batch = ChainBatch([model] * len(chains), chains)
    autotune.apply(batch, tune_w)
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
    default_shape = len(chains) == 1024 and args.sweeps == 500 and tuple(tune_w) == TRAFFIC_SHAPE
    roofline = roofline_of(launch_bytes, avg_kernel_ms, lds_meas)
    roofline.update({"traffic": TRAFFIC_DEFAULT_LAUNCH if default_shape else None, "hbm_copy_measured": hbm_meas,
                     "bytes_per_move": bytes_per_flip, "kernel_ms": avg_kernel_ms})
    out = {
        "metric": METRIC, "value": value, "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched SGC, binary FCC 10x10x10 (1000 sites), pairs<=4th + triplets<=2nd shell, "
                               "%d chains (64 T in 300-1500 K x 8 mu in [-0.2,0.2] eV x 2 seeds), %d sweeps "
                               "per step, one persistent launch" % (len(chains), args.sweeps),
                   "chains": len(chains), "sweeps_per_step": args.sweeps, "sites": settings.n_sites},
        "plain_moves_per_s": plain_value, "batched_over_plain": value / plain_value,
        # the batched workload counts SGC flips, the plain one canonical swaps (two flips each)
        "batched_over_plain_flips": value / (2.0 * plain_value),
        "roofline": roofline,
        "launch_shape": {"shape": autotune.shape_name(tune_w), "waves_per_chain": tune_w[0],
                         "moves_per_round": tune_w[1] if tune_w[2] != autotune.PATH_FIELDS else None,
                         "energy_path": "fields" if tune_w[2] == autotune.PATH_FIELDS else "evaluate",
                         "autotune_moves_per_s": {autotune.shape_name(k): v for k, v in tune_rates.items()}},
    }
    work = set(w.strip() for w in args.workloads.split(","))
    if "dropin" in work:
        out["dropin_updater"] = dropin_rate(settings, eci)
        if not args.no_cpu:
            ref1 = reference_sample("canonical", 3.0, 1)
            if ref1 is not None:
                out["dropin_updater"]["reference_moves_per_s"] = ref1["moves_per_s"]
        log("drop-in updater: %.3e moves/s (one launch per call: %.3e)" % (
            out["dropin_updater"]["moves_per_s"], out["dropin_updater"]["launch_per_call_moves_per_s"]))
    if "ternary" in work:
        out["ternary_uneven"] = ternary_uneven(lds_peak=lds_meas, cpu=not args.no_cpu)
        log("ternary uneven: %.3e moves/s (kernel %.1f ms)" % (out["ternary_uneven"]["value"],
                                                               out["ternary_uneven"]["kernel_ms"]))
    if "tempering" in work:
        out["parallel_tempering"] = tempering(cycles=args.pt_cycles, lds_peak=lds_meas, cpu=not args.no_cpu)
        log("tempering: %.3e moves/s (kernel %.1f ms)" % (out["parallel_tempering"]["value"],
                                                          out["parallel_tempering"]["kernel_ms"]))
    if "wang_landau" in work:
        out["wang_landau"] = wang_landau(model, settings)
        log("wang-landau: %.3e swaps/s (kernel %.1f ms); shared histogram %.3e swaps/s" % (
            out["wang_landau"]["value"], out["wang_landau"]["kernel_ms"],
            out["wang_landau"]["shared_histogram"]["value"]))
    if "serving" in work:
        out["serving"] = serving(n_jobs=args.serving_jobs)
        log("serving: %d jobs, p50 %.2f s, p99 %.2f s, %.3e moves/s" % (
            out["serving"]["jobs"], out["serving"]["latency_p50_s"], out["serving"]["latency_p99_s"],
            out["serving"]["throughput_moves_per_s"]))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(settings, eci, chains)
        log("cpu baseline: %.3e moves/s on %d threads" % (out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
