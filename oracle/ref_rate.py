"""Test infrastructure: times the reference's own CEUpdater (compiled from the
reference sources into oracle/_ref by ``make -C oracle ref``) on binary FCC
10x10x10 (pairs to the 4th shell, triplets to the 2nd), driven one move at a
time by the fed-stream sampler loop (cemc_amd.dropin.run_updater_chain, the
reference's Montecarlo._mc_step restated over the updater interface), the
way the reference's samplers call it.  Run with the interpreter the binding
is built for:

    /opt/conda/bin/python3.9 oracle/ref_rate.py KIND SECONDS PROCESSES

KIND "sgc": chain p flips at the p-th (temperature, chemical potential) of
the batched workload (64 temperatures geomspace(300, 1500) x 8 potentials
linspace(-0.2, 0.2)); "canonical": swaps at 800 K.  PROCESSES independent
chains run in parallel processes, each for about SECONDS.  Prints one JSON
object.  Only bench.py's CPU-baseline leg runs it.
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(HERE, "_ref"))


def _chain(args):
    kind, budget, p = args
    import numpy as np

    import cemc_ref
    from cemc_amd.dropin import run_updater_chain
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_occupation, seeded_ecis
    from oracle.brute_force import brute_force_cf
    # no equivalent-site lists: the reference would import ase.clease for
    # them (cluster.cpp:77-81); a binary model's decorations are all 0 anyway
    s = FCCSettings(10, ["A", "B"], {2: 4, 3: 2}, symmetrize_decorations=False)
    names = s.cf_names()
    eci = seeded_ecis(s, 2024, 0.02)
    if kind == "sgc":
        T = float(np.geomspace(300.0, 1500.0, 64)[p % 64])
        eci["c1_0"] -= float(np.linspace(-0.2, 0.2, 8)[(p // 64) % 8])
        move_kind, occ = 1, random_occupation(s.n_sites, 2, 1000 + p)
    else:
        T, move_kind, occ = 800.0, 0, random_occupation(s.n_sites, 2, 1)
    cf0 = brute_force_cf(s, occ, names)
    s.trans_matrix = s.trans_matrix.to_list()
    atoms = SimpleAtoms([s.species[v] for v in occ])
    upd = cemc_ref.RefUpdater(atoms, s, dict(cf0), dict(eci))
    # calibrate, then time a bounded sample of the same chain's stream
    n = 1000
    t0 = time.perf_counter()
    run_updater_chain(upd, atoms, s.species, move_kind, 77, p, n, T)
    dt = time.perf_counter() - t0
    n = max(1000, int(n * budget / max(dt, 1e-3)))
    t0 = time.perf_counter()
    run_updater_chain(upd, atoms, s.species, move_kind, 77, p, n, T)
    return n, time.perf_counter() - t0


def main():
    kind = sys.argv[1] if len(sys.argv) > 1 else "sgc"
    budget = float(sys.argv[2]) if len(sys.argv) > 2 else 5.0
    procs = int(sys.argv[3]) if len(sys.argv) > 3 else 1
    t0 = time.perf_counter()
    if procs > 1:
        with mp.get_context("fork").Pool(procs) as pool:
            res = pool.map(_chain, [(kind, budget, p) for p in range(procs)])
    else:
        res = [_chain((kind, budget, 0))]
    wall = time.perf_counter() - t0
    print(json.dumps({"kind": kind, "processes": procs, "moves": sum(n for n, _ in res),
                      "moves_per_s": sum(n for n, _ in res) / max(t for _, t in res),
                      "moves_per_s_per_process": [n / t for n, t in res], "wall_s": wall}))


if __name__ == "__main__":
    main()
