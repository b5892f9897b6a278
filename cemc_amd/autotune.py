"""Launch-shape autotuner: picks the number of wavefronts per chain (1, 2, 4
or 8 cooperating waves) for a batch by timing each candidate on a scratch
copy of the batch, so the real chains' trajectories are not touched.

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


def tune(models, chains, candidates=(1, 2, 4, 8), sweeps=2.0, repeats=2, cache_file=None, log=None):
    """Returns (best waves, {waves: moves/s}) for the batch (models, chains).

    Each candidate runs ``sweeps`` sweeps of every chain on a scratch batch
    built from the same inputs, ``repeats`` times; the best rate counts."""
    sig = signature(models, chains)
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if sig in _cache:
        entry = _cache[sig]
        return int(entry["best"]), {int(k): v for k, v in entry["rates"].items()}
    scratch = ChainBatch(models, chains)
    moves = np.array([max(64, int(sweeps * m.n_sites)) for m in models], dtype=np.int64)
    rates = {}
    seen = set()
    for w in candidates:
        scratch.set_waves(w)
        shape = tuple(sorted({scratch.launch_shape(c) for c in range(scratch.n_chains)}))
        if shape in seen:  # the variant cannot take this many waves: same launch as before
            continue
        seen.add(shape)
        scratch.run(moves // 4)  # first launch of a variant loads its code
        best = 0.0
        for _ in range(repeats):
            ms = scratch.run_timed(moves)
            best = max(best, float(moves.sum()) / (ms * 1e-3))
        rates[w] = best
        if log:
            log("autotune %s: waves %d -> %.3e moves/s" % (sig, w, best))
    scratch.close()
    best_w = max(rates, key=rates.get)
    _cache[sig] = {"best": best_w, "rates": {str(k): v for k, v in rates.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
    return best_w, rates


def tuned_batch(models, chains, **kw):
    """A ChainBatch with the autotuned waves per chain applied."""
    best, _ = tune(models, chains, **kw)
    batch = ChainBatch(models, chains)
    batch.set_waves(best)
    return batch
