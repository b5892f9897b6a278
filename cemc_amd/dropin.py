"""The reference's per-move sampler loop over the updater interface, on the
fed Philox move stream: the drop-in path that a caller of
``cemc_amd.updater.PyCEUpdater`` (or of the reference's own ``PyCEUpdater``)
runs when it drives one move at a time as ``Montecarlo._mc_step`` does.
The same function drives the reference's compiled updater when the golden
fixtures are written, and this package's updater in the parity tests and the
bench, so both consume an identical move stream.
"""
from __future__ import annotations

import numpy as np

from .lattice import KB_EV

MASK64 = (1 << 64) - 1


def philox_blocks(seed, chain_uid, move_begin, n_moves):
    """Consecutive Philox4x64-10 blocks for counters (t, 0, 0, 0)."""
    start = (move_begin - 1) & ((1 << 256) - 1)  # numpy increments before use
    ctr = np.array([(start >> (64 * i)) & MASK64 for i in range(4)], dtype=np.uint64)
    bg = np.random.Philox(key=np.array([seed, chain_uid], dtype=np.uint64), counter=ctr)
    return bg.random_raw(4 * n_moves).reshape(n_moves, 4)


def bounded64(w, n):
    return (int(w) * int(n)) >> 64


def bounded32(w, n):
    return (int(w) * int(n)) >> 32


def unit_double(w):
    return float(int(w) >> 11) * 2.0 ** -53


def run_updater_chain(upd, atoms, species, kind, seed, chain_uid, n_moves, temperature,
                      trace_every=0, cf_every=0, names=(), background=()):
    """The reference's sampler step over an updater with the PyCEUpdater
    interface (calculate, clear_history, undo_changes, get_energy, get_cf),
    one call per move as Montecarlo._mc_step makes them through CE
    (cemc/mcmc/montecarlo.py:970-1038, _accept :910-956, _get_trial_move
    :890-908; cemc/mcmc/sgc_montecarlo.py:62-76), with the proposal
    randomness of the fed stream (move t: Philox4x64-10 block key (seed,
    chain_uid), counter (t, 0, 0, 0); w0 hi32 -> first species, lo32 ->
    offset of the second; w1 -> slot in the first species' site list (SGC:
    movable site); w2 -> slot in the second species' list; w3 -> the uniform,
    always drawn).  `atoms` is the updater's atoms object (kept in step by
    the updater).  Returns (accept bits LSB first, energy every trace_every
    moves, CFs every cf_every moves, final species ids, accepted moves)."""
    n_s = len(species)
    sym = [a.symbol for a in atoms]
    sid = {s: i for i, s in enumerate(species)}
    occ = [sid[s] for s in sym]
    n_sites = len(occ)
    bkg = set(int(i) for i in background)
    movable = [i for i in range(n_sites) if i not in bkg]
    if kind == 0:
        present = [s for s in range(n_s) if any(occ[i] == s for i in movable)]
        tracker = {s: [i for i in movable if occ[i] == s] for s in present}
    blocks = philox_blocks(seed, chain_uid, 0, n_moves)
    bits = np.zeros((n_moves + 7) // 8, dtype=np.uint8)
    e_trace, cf_ck = [], []
    e_cur = upd.get_energy()
    kT = temperature * KB_EV
    n_acc = 0
    for t in range(n_moves):
        w0, w1, w2, w3 = (int(v) for v in blocks[t])
        if kind == 0:
            ia = bounded32(w0 >> 32, len(present))
            ib = (ia + 1 + bounded32(w0 & 0xFFFFFFFF, len(present) - 1)) % len(present)
            a, b = present[ia], present[ib]
            sa = bounded64(w1, len(tracker[a]))
            sb = bounded64(w2, len(tracker[b]))
            pa, pb = tracker[a][sa], tracker[b][sb]
            changes = [(pa, species[a], species[b]), (pb, species[b], species[a])]
        else:
            site = movable[bounded64(w1, len(movable))]
            old = sid[atoms[site].symbol]
            new = (old + 1 + bounded32(w0 & 0xFFFFFFFF, n_s - 1)) % n_s
            changes = [(site, species[old], species[new])]
        u = unit_double(w3)
        e_new = upd.calculate(changes)
        if e_new < e_cur:
            accept = True
        else:
            energy_diff = e_new - e_cur
            accept = u <= np.exp(-energy_diff / kT)
        if accept:
            upd.clear_history()
            e_cur = e_new
            n_acc += 1
            bits[t >> 3] |= np.uint8(1 << (t & 7))
            if kind == 0:
                tracker[a][sa] = pb
                tracker[b][sb] = pa
        else:
            upd.undo_changes()
        if trace_every and (t + 1) % trace_every == 0:
            e_trace.append(e_cur)
        if cf_every and (t + 1) % cf_every == 0:
            cf = upd.get_cf()
            cf_ck.append([cf[n] for n in names])
    final = np.array([sid[a.symbol] for a in atoms], dtype=np.int8)
    return bits, np.array(e_trace), np.array(cf_ck), final, n_acc
