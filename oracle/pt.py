"""Test infrastructure: CPU restatement of parallel tempering on the fed
stream (only tests use it).

Follows the reference's exchange step (cemc/mcmc/parallel_tempering.py:
138-144 acceptance, 146-151 exchange, 153-175 pair lists, 177-191 cycle)
with the random numbers of the stream that include/cemc_gpu.h fixes: the
direction of cycle c from Philox block (c, 0, 0, 0), the uniform of pair k
from block (c, 1 + k, 0, 0), both with key (seed, uid).  Exchanging
temperatures between two lattices is the same move as the reference's
exchange of configurations between two temperatures.

Pinned (tests/test_pt_oracle.py) by what the mathematics fixes, not by a
golden vector (the reference's sampler needs ase, absent here): with the
lattices frozen the slot permutation samples exp(-sum_t E_sigma(t)/kT_t)
exactly; with sweeps each slot samples its own Boltzmann law and each
pair's exchange acceptance equals the exact expectation of
min(1, exp(d(1/kT) dE)) under the product of the two laws; the pairs tried
follow the reference's up/down lists.
"""
from __future__ import annotations

import math

import numpy as np

from .oracle import philox4x64

KB_EV = 8.617330337217213e-05


def accept_probability(E1, E2, T1, T2):
    """ParallelTempering._accept_probability (parallel_tempering.py:138-144):
    exp((1/kT1 - 1/kT2) (E1 - E2)), not capped at 1."""
    b1 = 1.0 / (KB_EV * T1)
    b2 = 1.0 / (KB_EV * T2)
    return math.exp((b1 - b2) * (E1 - E2))


def exchange_pairs(R, up):
    """_perform_exchange_move's pair lists (parallel_tempering.py:153-160):
    up = (0,1), (2,3), ...; down = (R-1,R-2), (R-3,R-4), ..."""
    if up:
        return [(i, i + 1) for i in range(0, R - 1, 2)]
    return [(i, i - 1) for i in range(R - 1, 0, -2)]


def run_pt(updaters, kind, chain_seeds, chain_uids, temperatures, n_cycles, moves_per_cycle, seed, uid,
           record=None, accept=accept_probability, pairs=exchange_pairs):
    """updaters[c]: OracleUpdater of chain c, initially at slot c with
    temperatures[c].  Returns (chain_of_slot, n_accept per pair, per-slot
    lists of energies sampled after every move).  `record` (a dict) receives
    "n_try" per pair and "history", the chain of each slot after every
    cycle.  `accept` and `pairs` default to the reference's rules (tests pass
    planted variants to show the checks catch them)."""
    R = len(updaters)
    cos = list(range(R))
    soc = list(range(R))
    n_acc = [0] * (R - 1)
    n_try = [0] * (R - 1)
    history = []
    samples = [[] for _ in range(R)]
    for c in range(n_cycles):
        for ch in range(R):
            slot = soc[ch]
            if moves_per_cycle > 0:
                _, trace, _ = updaters[ch].run(kind, chain_seeds[ch], chain_uids[ch], moves_per_cycle,
                                               temperatures[slot], move_begin=c * moves_per_cycle, trace_every=1)
                samples[slot].extend(trace.tolist())
        up = philox4x64((c, 0, 0, 0), (seed, uid))[0] >> 63
        for k, (a, b) in enumerate(pairs(R, up)):
            ca, cb = cos[a], cos[b]
            ea, eb = updaters[ca].get_energy(), updaters[cb].get_energy()
            p = accept(ea, eb, temperatures[a], temperatures[b])
            u = float(philox4x64((c, 1 + k, 0, 0), (seed, uid))[3] >> 11) * 2.0 ** -53
            n_try[min(a, b)] += 1
            if u < p:
                n_acc[min(a, b)] += 1
                cos[a], cos[b] = cb, ca
                soc[cb], soc[ca] = a, b
        if record is not None:
            history.append(tuple(cos))
    if record is not None:
        record["n_try"] = np.array(n_try)
        record["history"] = history
    return np.array(cos), np.array(n_acc), [np.array(s) for s in samples]
