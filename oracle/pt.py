"""Test infrastructure: CPU restatement of parallel tempering on the fed
stream (only tests use it).

Follows the reference's exchange step (cemc/mcmc/parallel_tempering.py:
138-144 acceptance, 146-151 exchange, 153-175 pair lists, 177-191 cycle)
with the random numbers of the stream that include/cemc_gpu.h fixes: the
direction of cycle c from Philox block (c, 0, 0, 0), the uniform of pair k
from block (c, 1 + k, 0, 0), both with key (seed, uid).  Exchanging
temperatures between two lattices is the same move as the reference's
exchange of configurations between two temperatures.
"""
from __future__ import annotations

import math

import numpy as np

from .oracle import philox4x64

KB_EV = 8.617330337217213e-05


def run_pt(updaters, kind, chain_seeds, chain_uids, temperatures, n_cycles, moves_per_cycle, seed, uid):
    """updaters[c]: OracleUpdater of chain c, initially at slot c with
    temperatures[c].  Returns (chain_of_slot, n_accept per pair, per-slot
    lists of energies sampled after every move)."""
    R = len(updaters)
    cos = list(range(R))
    soc = list(range(R))
    kT = [T * KB_EV for T in temperatures]
    n_acc = [0] * (R - 1)
    samples = [[] for _ in range(R)]
    for c in range(n_cycles):
        for ch in range(R):
            slot = soc[ch]
            _, trace, _ = updaters[ch].run(kind, chain_seeds[ch], chain_uids[ch], moves_per_cycle,
                                           temperatures[slot], move_begin=c * moves_per_cycle, trace_every=1)
            samples[slot].extend(trace.tolist())
        up = philox4x64((c, 0, 0, 0), (seed, uid))[0] >> 63
        for k in range(R // 2):
            a = 2 * k if up else R - 1 - 2 * k
            b = a + 1 if up else a - 1
            ca, cb = cos[a], cos[b]
            ea, eb = updaters[ca].get_energy(), updaters[cb].get_energy()
            p = math.exp((1.0 / kT[a] - 1.0 / kT[b]) * (ea - eb))
            u = float(philox4x64((c, 1 + k, 0, 0), (seed, uid))[3] >> 11) * 2.0 ** -53
            if u < p:
                n_acc[min(a, b)] += 1
                cos[a], cos[b] = cb, ca
                soc[cb], soc[ca] = a, b
    return np.array(cos), np.array(n_acc), [np.array(s) for s in samples]
