"""Launch-shape autotuner: picks, for a batch, the energy path (per-site
fields updated on accepted moves, one wave per chain; or tables summed per
trial move) and, for the latter, the number of wavefronts per chain (1, 2, 4
or 8 cooperating waves) and the moves evaluated per round (1, 8, or 8 per
wave: wide rounds, where each wave evaluates its own eight), by timing
each candidate on a scratch copy of the batch, so the real chains'
trajectories are not touched.

The kernel variant (records per lane, energy tables in LDS or HBM) follows
from the shape and the model.  More waves split each move's sub-clusters
further (shorter dependent chain per move, one barrier per move or round
more), fewer waves leave more CUs to other chains; fields make rejected
moves cheap and accepted ones dearer.  A shape is (waves, moves per round,
path); its name is "fields" or "WxK".  Results are cached per workload
signature in memory and, optionally, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .batch import PATH_EVALUATE, PATH_FIELDS, ChainBatch

_cache = {}


def signature(models, chains):
    """Workload signature: per distinct (model size, species, sub-clusters,
    move kind) the number of chains."""
    sig = {}
    for m, ch in zip(models, chains):
        key = "%d/%d/%d/%d" % (m.n_sites, m.n_species, m.n_subclusters, int(ch.get("kind", 0)))
        sig[key] = sig.get(key, 0) + 1
    return ";".join("%s:%d" % kv for kv in sorted(sig.items()))


def shape_name(shape):
    return "fields" if shape[2] == PATH_FIELDS else "%dx%d" % (shape[0], shape[1])


def parse_shape(name):
    """"fields" or "WxK" -> (waves, moves per round, path)."""
    if name == "fields":
        return (1, 1, PATH_FIELDS)
    w, k = (int(v) for v in name.split("x"))
    return (w, k, PATH_EVALUATE)


def tune(models, chains, candidates=(1, 2, 4, 8), per_round=(1, 8), fields=True, sweeps=2.0, repeats=2,
         cache_file=None, log=None, wide=True):
    """Returns (best shape, {shape: moves/s}) for the batch (models, chains).

    Each candidate shape runs ``sweeps`` sweeps of every chain on a scratch
    batch built from the same inputs, ``repeats`` times; the best rate counts.
    The fields path is a candidate when some chain of the batch can take it."""
    sig = signature(models, chains)
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if sig in _cache:
        entry = _cache[sig]
        return tuple(entry["best"]), {parse_shape(k): r for k, r in entry["rates"].items()}
    scratch = ChainBatch(models, chains)
    moves = np.array([max(64, int(sweeps * m.n_sites)) for m in models], dtype=np.int64)
    shapes = []
    if fields:
        scratch.set_energy_path(PATH_FIELDS)
        if any(scratch.energy_path(c) == PATH_FIELDS for c in range(scratch.n_chains)):
            shapes.append((1, 1, PATH_FIELDS))
    shapes += [(w, k, PATH_EVALUATE) for k in per_round for w in candidates]
    if wide:
        # wide rounds: each of w waves evaluates its own eight moves
        shapes += [(w, 8 * w, PATH_EVALUATE) for w in candidates if w > 1]
    rates = {}
    seen = set()
    for shape in shapes:
        apply(scratch, shape)
        launch = (shape[1], shape[2], tuple(sorted({scratch.launch_shape(c) for c in range(scratch.n_chains)})))
        if launch in seen:  # the variant cannot take this shape: same launch as before
            continue
        seen.add(launch)
        scratch.run(moves // 4)  # first launch of a variant loads its code
        best = 0.0
        for _ in range(repeats):
            ms = scratch.run_timed(moves)
            best = max(best, float(moves.sum()) / (ms * 1e-3))
        rates[shape] = best
        if log:
            log("autotune %s: %s -> %.3e moves/s" % (sig, shape_name(shape), best))
    scratch.close()
    best = max(rates, key=rates.get)
    _cache[sig] = {"best": list(best), "rates": {shape_name(k): v for k, v in rates.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
    return best, rates


def apply(batch, shape):
    """Set a tuned (waves, moves per round, path) on a batch."""
    batch.set_energy_path(shape[2])
    batch.set_moves_per_round(shape[1])
    batch.set_waves(shape[0])


def tuned_batch(models, chains, **kw):
    """A ChainBatch with the autotuned launch shape applied."""
    best, _ = tune(models, chains, **kw)
    batch = ChainBatch(models, chains)
    apply(batch, best)
    return batch
