"""Launch-shape autotuner: picks the number of wavefronts per chain (1, 2, 4
or 8 cooperating waves) and the moves evaluated per round (1 or 8) for a
batch by timing each candidate on a scratch copy of the batch, so the real
chains' trajectories are not touched.

The kernel variant (records per lane, energy tables in LDS or HBM) follows
from the waves and the model; the waves are the free parameter: more waves
split each move's sub-clusters further (shorter dependent chain per move,
one barrier per move more), fewer waves leave more CUs to other chains.
Results are cached per workload signature in memory and, optionally, in a
JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .batch import ChainBatch

_cache = {}


def signature(models, chains):
    """Workload signature: per distinct (model size, species, sub-clusters,
    move kind) the number of chains."""
    sig = {}
    for m, ch in zip(models, chains):
        key = "%d/%d/%d/%d" % (m.n_sites, m.n_species, m.n_subclusters, int(ch.get("kind", 0)))
        sig[key] = sig.get(key, 0) + 1
    return ";".join("%s:%d" % kv for kv in sorted(sig.items()))


def tune(models, chains, candidates=(1, 2, 4, 8), per_round=(1, 8), sweeps=2.0, repeats=2, cache_file=None,
         log=None):
    """Returns ((waves, moves per round) of the best shape, {shape: moves/s})
    for the batch (models, chains).

    Each candidate shape runs ``sweeps`` sweeps of every chain on a scratch
    batch built from the same inputs, ``repeats`` times; the best rate counts."""
    sig = signature(models, chains)
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if sig in _cache:
        entry = _cache[sig]
        return tuple(entry["best"]), {tuple(int(v) for v in k.split("x")): r for k, r in entry["rates"].items()}
    scratch = ChainBatch(models, chains)
    moves = np.array([max(64, int(sweeps * m.n_sites)) for m in models], dtype=np.int64)
    rates = {}
    seen = set()
    for k in per_round:
        scratch.set_moves_per_round(k)
        for w in candidates:
            scratch.set_waves(w)
            shape = (k, tuple(sorted({scratch.launch_shape(c) for c in range(scratch.n_chains)})))
            if shape in seen:  # the variant cannot take this shape: same launch as before
                continue
            seen.add(shape)
            scratch.run(moves // 4)  # first launch of a variant loads its code
            best = 0.0
            for _ in range(repeats):
                ms = scratch.run_timed(moves)
                best = max(best, float(moves.sum()) / (ms * 1e-3))
            rates[(w, k)] = best
            if log:
                log("autotune %s: waves %d, moves per round %d -> %.3e moves/s" % (sig, w, k, best))
    scratch.close()
    best = max(rates, key=rates.get)
    _cache[sig] = {"best": list(best), "rates": {"%dx%d" % kk: v for kk, v in rates.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
    return best, rates


def apply(batch, shape):
    """Set a tuned (waves, moves per round) on a batch."""
    batch.set_moves_per_round(shape[1])
    batch.set_waves(shape[0])


def tuned_batch(models, chains, **kw):
    """A ChainBatch with the autotuned launch shape applied."""
    best, _ = tune(models, chains, **kw)
    batch = ChainBatch(models, chains)
    apply(batch, best)
    return batch
