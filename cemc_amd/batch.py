"""Batched Markov chains resident on the GPU.

Each chain is one lattice of a ``CEModel`` with its own temperature, move
kind (canonical swaps or semi-grand-canonical flips), chemical potentials
and Philox stream key.  ``run`` advances every chain by its own number of
trial moves in one persistent launch (one wavefront per chain).  Per move the
kernel restates the reference sampler step (cemc/mcmc/montecarlo.py:970-1038)
with the proposal distributions of montecarlo.py:890-908 and
sgc_montecarlo.py:62-76 drawn from the fed stream: move t of a chain uses
Philox4x64-10 with key (seed, uid) and counter (t, 0, 0, 0).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi

CANONICAL = _capi.MOVE_CANONICAL
SGC = _capi.MOVE_SGC
PATH_AUTO, PATH_EVALUATE, PATH_FIELDS = _capi.PATH_AUTO, _capi.PATH_EVALUATE, _capi.PATH_FIELDS


def chem_pot_vector(model, chem_pot):
    """Chemical potentials keyed by singlet name -> array indexed by decoration."""
    vec = np.zeros(max(model.n_species - 1, 1))
    if chem_pot is None:
        return vec
    if isinstance(chem_pot, dict):
        for name, val in chem_pot.items():
            if not name.startswith("c1_") or name not in model.names:
                raise ValueError("A chemical potential that is currently not tracked is added: %s" % name)
            vec[int(name.rsplit("_", 1)[1])] = float(val)
        return vec
    arr = np.asarray(chem_pot, dtype=np.float64)
    vec[:len(arr)] = arr
    return vec


def _fill_desc(d, m, ch, keep):
    occ, op = _capi.i8(ch["occupation"])
    if occ.shape[0] != m.n_sites:
        raise ValueError("occupation length does not match the model")
    mu, mp = _capi.f64(chem_pot_vector(m, ch.get("chem_pot")))
    keep += [occ, mu]
    d.kind = int(ch.get("kind", CANONICAL))
    d.temperature = float(ch["temperature"])
    d.chem_pot = mp
    d.seed = int(ch.get("seed", 0)) & 0xFFFFFFFFFFFFFFFF
    d.uid = int(ch.get("uid", 0)) & 0xFFFFFFFFFFFFFFFF
    d.occupation = op
    d.first_move = int(ch.get("first_move", 0)) & 0xFFFFFFFFFFFFFFFF


class ChainBatch:
    def __init__(self, models, chains):
        """models: one CEModel per chain; chains: dicts with keys kind,
        temperature, occupation, seed, uid and optionally chem_pot and
        first_move (Philox counter of the first move, default 0)."""
        lib = _capi.load()
        n = len(chains)
        if n == 0 or len(models) != n:
            raise ValueError("one model per chain is required")
        self.models = list(models)
        self._lib = lib
        descs = (_capi.ChainDesc * n)()
        handles = (ctypes.c_void_p * n)()
        keep = []
        self.n_sites = np.zeros(n, dtype=np.int64)
        self.kinds = np.zeros(n, dtype=np.int32)
        self.temperatures = np.zeros(n)
        for c, (m, ch) in enumerate(zip(models, chains)):
            d = descs[c]
            _fill_desc(d, m, dict(ch, uid=ch.get("uid", c)), keep)
            handles[c] = m.handle
            self.n_sites[c] = m.n_sites
            self.kinds[c] = d.kind
            self.temperatures[c] = d.temperature
        h = ctypes.c_void_p()
        _capi.check(lib.cemc_batch_create(n, handles, descs, ctypes.byref(h)))
        self.handle = h
        self.n_chains = n

    def _moves(self, n_moves):
        if np.isscalar(n_moves):
            return None, int(n_moves)
        a = np.ascontiguousarray(n_moves, dtype=np.int64)
        if a.shape[0] != self.n_chains:
            raise ValueError("one move count per chain is required")
        return a, 0

    def run(self, n_moves):
        """Advance each chain by n_moves (int or one value per chain)."""
        a, scalar = self._moves(n_moves)
        ptr = a.ctypes.data_as(_capi._i64p) if a is not None else None
        _capi.check(self._lib.cemc_batch_run(self.handle, ptr, scalar))

    def run_timed(self, n_moves):
        """Like run; returns the launch's device time in milliseconds (HIP events)."""
        a, scalar = self._moves(n_moves)
        ptr = a.ctypes.data_as(_capi._i64p) if a is not None else None
        ms = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_batch_run_timed(self.handle, ptr, scalar, ctypes.byref(ms)))
        return ms.value

    def run_sliced(self, n_moves, max_ms):
        """Like run_timed, but every chain stops at the first 64-move block
        that starts more than max_ms after its own start on the device;
        returns the device time in ms.  stats()["moves"] tells how far each
        chain got; its stream continues there on the next run."""
        a, scalar = self._moves(n_moves)
        ptr = a.ctypes.data_as(_capi._i64p) if a is not None else None
        ms = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_batch_run_sliced(self.handle, ptr, scalar, float(max_ms), ctypes.byref(ms)))
        return ms.value

    def run_logged(self, n_moves):
        """Like run; returns per-chain accept bits (uint8, LSB first, one bit per move)."""
        a, scalar = self._moves(n_moves)
        ptr = a.ctypes.data_as(_capi._i64p) if a is not None else None
        most = int(a.max()) if a is not None else scalar
        words = max((most + 31) // 32, 1)
        log = np.zeros((self.n_chains, words), dtype=np.uint32)
        _capi.check(self._lib.cemc_batch_run_logged(self.handle, ptr, scalar,
                                                    log.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), words))
        return log.view(np.uint8).reshape(self.n_chains, -1)

    def run_sweeps(self, n_sweeps):
        """One sweep = n_sites trial moves of that chain."""
        return self.run(np.asarray(n_sweeps, dtype=np.int64) * self.n_sites)

    def energies(self):
        out = np.zeros(self.n_chains)
        _capi.check(self._lib.cemc_batch_get_energy(self.handle, out.ctypes.data_as(_capi._dp)))
        return out

    def occupation(self, chain):
        out = np.zeros(int(self.n_sites[chain]), dtype=np.int8)
        _capi.check(self._lib.cemc_batch_get_occupation(self.handle, int(chain), out.ctypes.data_as(_capi._i8p)))
        return out

    def stats(self):
        n = self.n_chains
        moves = np.zeros(n, dtype=np.int64)
        acc = np.zeros(n, dtype=np.int64)
        se = np.zeros(n)
        se2 = np.zeros(n)
        _capi.check(self._lib.cemc_batch_get_stats(self.handle, moves.ctypes.data_as(_capi._i64p),
                                                   acc.ctypes.data_as(_capi._i64p), se.ctypes.data_as(_capi._dp),
                                                   se2.ctypes.data_as(_capi._dp)))
        return {"moves": moves, "accepted": acc, "sum_e": se, "sum_e2": se2}

    def reset_stats(self, chain=-1):
        """Zero the sample sums of one chain (or all, chain=-1); the current
        energy becomes their reference."""
        _capi.check(self._lib.cemc_batch_reset_stats(self.handle, int(chain)))

    def thermo(self, chain):
        """Sample averages of one chain since its last reset: n_samples,
        energy (current), mean_energy, var_energy, singlet, var_singlet,
        cov_singlet_energy (arrays over decorations)."""
        t = _capi.Thermo()
        _capi.check(self._lib.cemc_batch_get_thermo(self.handle, int(chain), ctypes.byref(t)))
        b = int(t.n_singlets)
        return {"n_samples": int(t.n_samples), "energy": t.energy, "mean_energy": t.mean_energy,
                "var_energy": t.var_energy, "singlet": np.array(t.singlet[:b]),
                "var_singlet": np.array(t.var_singlet[:b]),
                "cov_singlet_energy": np.array(t.cov_singlet_energy[:b])}

    def cf(self, chain):
        """Correlation functions of one chain, computed from scratch on the GPU,
        and the energy they give."""
        m = self.models[chain]
        out = np.zeros(m.n_cf)
        e = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_batch_get_cf(self.handle, int(chain), out.ctypes.data_as(_capi._dp),
                                                ctypes.byref(e)))
        return out, e.value

    def set_waves(self, waves):
        """Force the waves per chain (1, 2, 4, 8) where the kernel variant
        allows it; 0 restores the automatic choice."""
        _capi.check(self._lib.cemc_batch_set_waves(self.handle, int(waves)))

    def set_moves_per_round(self, moves):
        """0 (default): eight consecutive moves evaluated together wherever
        the chain's kernel variant allows it and the moves of a round seldom
        conflict (interacting sites, times 4 for swaps, at most 12% of the
        lattice), else one at a time; 1: every move on
        its own; 8: rounds of eight where allowed; 16, 32, 64: wide rounds
        on 2, 4, 8 waves per chain, each wave evaluating its own eight moves
        of the round (periodic lattices whose records fit one wave; other
        chains take rounds of eight); wide rounds pay on large lattices at
        low acceptance, so the autotuner times them rather than choosing
        them by default.  Same results bit for bit."""
        _capi.check(self._lib.cemc_batch_set_moves_per_round(self.handle, int(moves)))

    def set_energy_path(self, path):
        """PATH_AUTO (default) and PATH_EVALUATE: every trial move sums its
        sub-clusters' table entries; PATH_FIELDS: per-site fields kept in LDS
        and updated on accepted moves, where the chain allows it (fast for
        single chains at low acceptance).  Same accept sequences."""
        _capi.check(self._lib.cemc_batch_set_energy_path(self.handle, int(path)))

    def energy_path(self, chain):
        """PATH_EVALUATE or PATH_FIELDS: the path chain `chain` runs on."""
        v = ctypes.c_int32()
        _capi.check(self._lib.cemc_batch_get_energy_path(self.handle, int(chain), ctypes.byref(v)))
        return v.value

    def launch_shape(self, chain):
        """(waves per chain, records per lane in registers, tables in LDS)."""
        w, r, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _capi.check(self._lib.cemc_batch_get_launch_shape(self.handle, int(chain), ctypes.byref(w),
                                                          ctypes.byref(r), ctypes.byref(g)))
        return w.value, r.value, bool(g.value)

    def set_temperature(self, chain, temperature):
        _capi.check(self._lib.cemc_batch_set_temperature(self.handle, int(chain), float(temperature)))
        self.temperatures[chain] = float(temperature)

    # ---------------------------------------------- constraints and bias
    def set_hooks(self, chain, fixed_species=0, max_attempts=10000, allowed=None, pair=None, pair_neighbours=None,
                  bias_species=-1, bias_table=None, recycle_waste=False):
        """Constraints and bias of chain `chain` (cemc_batch_set_hooks).
        allowed: uint8 mask per site of the species it may hold; pair: (e1, e2)
        species ids with pair_neighbours an (n_sites, members) int array;
        bias_table: n_sites + 1 energies over the count of bias_species;
        recycle_waste: heat-bath acceptance and recycled samples."""
        d = _capi.HooksDesc()
        d.fixed_species = int(fixed_species)
        d.max_attempts = int(max_attempts)
        d.recycle_waste = int(bool(recycle_waste))
        keep = []
        n = int(self.n_sites[chain])
        if allowed is not None:
            a = np.ascontiguousarray(allowed, dtype=np.uint8)
            if a.shape != (n,):
                raise ValueError("allowed needs one mask per site")
            keep.append(a)
            d.allowed = a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.pair_species[0], d.pair_species[1] = (-1, -1) if pair is None else (int(pair[0]), int(pair[1]))
        if pair is not None:
            nb = np.ascontiguousarray(pair_neighbours, dtype=np.int32)
            if nb.ndim != 2 or nb.shape[0] != n:
                raise ValueError("pair_neighbours needs one row per site")
            keep.append(nb)
            d.pair_members = nb.shape[1]
            d.pair_neighbours = nb.ctypes.data_as(_capi._i32p)
        d.bias_species = int(bias_species)
        if bias_species >= 0:
            t = np.ascontiguousarray(bias_table, dtype=np.float64)
            if t.shape != (n + 1,):
                raise ValueError("bias_table needs n_sites + 1 entries")
            keep.append(t)
            d.bias_table = t.ctypes.data_as(_capi._dp)
        _capi.check(self._lib.cemc_batch_set_hooks(self.handle, int(chain), ctypes.byref(d)))

    def clear_hooks(self, chain):
        _capi.check(self._lib.cemc_batch_set_hooks(self.handle, int(chain), None))

    def hooks_state(self, chain):
        """(constrained pair count, biased species count, bias energy)."""
        pc, bc, be = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
        _capi.check(self._lib.cemc_batch_get_hooks_state(self.handle, int(chain), ctypes.byref(pc), ctypes.byref(bc),
                                                         ctypes.byref(be)))
        return pc.value, bc.value, be.value

    # ------------------------------------------------------- Wang-Landau
    def wl_set(self, chain, n_bins, e_min, e_max, f, f_min=1e-6, flatness=0.8, current_bin=0, is_first=True,
               converged=False, logdos=None, hist=None, known=None):
        """Give canonical or semi-grand-canonical chain `chain` a Wang-Landau walker
        (cemc_batch_wl_set)."""
        d = _capi.WLDesc(int(n_bins), int(current_bin), int(bool(is_first)), int(bool(converged)), float(e_min),
                         float(e_max), float(f), float(f_min), float(flatness))
        arrs = []
        for a, dt, ct in ((logdos, np.float64, ctypes.c_double), (hist, np.int32, ctypes.c_int32),
                          (known, np.uint8, ctypes.c_uint8)):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=dt)
            if a.shape != (int(n_bins),):
                raise ValueError("histogram arrays need n_bins entries")
            arrs.append(a)
        ptrs = [a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None
                for a, ct in zip(arrs, (ctypes.c_double, ctypes.c_int32, ctypes.c_uint8))]
        _capi.check(self._lib.cemc_batch_wl_set(self.handle, int(chain), ctypes.byref(d), *ptrs))

    def wl_get(self, chain):
        """Walker of chain `chain`: dict with the descriptor fields, the
        arrays logdos, hist, known and the counters."""
        d = _capi.WLDesc()
        _capi.check(self._lib.cemc_batch_wl_get(self.handle, int(chain), ctypes.byref(d), None, None, None, None))
        n = d.n_bins
        logdos = np.zeros(n)
        hist = np.zeros(n, dtype=np.int32)
        known = np.zeros(n, dtype=np.uint8)
        cnt = np.zeros(5, dtype=np.int64)
        _capi.check(self._lib.cemc_batch_wl_get(self.handle, int(chain), None, logdos.ctypes.data_as(_capi._dp),
                                                hist.ctypes.data_as(_capi._i32p),
                                                known.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                cnt.ctypes.data_as(_capi._i64p)))
        out = {k: getattr(d, k) for k, _ in _capi.WLDesc._fields_}
        out.update(logdos=logdos, hist=hist, known=known, n_moves=int(cnt[0]), n_accept=int(cnt[1]),
                   n_outside=int(cnt[2]), n_self=int(cnt[3]), n_flat=int(cnt[4]))
        return out

    def wl_set_scheme(self, chain, inverse_time=False, adaptive=False, min_width=10):
        """Schemes of walker `chain` (cemc_batch_wl_set_scheme): the
        inverse-time law and adaptive energy windows; starts them afresh."""
        sc = _capi.WLScheme()
        sc.inverse_time = int(bool(inverse_time))
        sc.adaptive = int(bool(adaptive))
        sc.min_width = int(min_width)
        _capi.check(self._lib.cemc_batch_wl_set_scheme(self.handle, int(chain), ctypes.byref(sc)))

    def wl_scheme(self, chain):
        """Scheme state of walker `chain`: dict of the cemc_wl_scheme fields
        plus arrays edges, edge_n, edge_vals (n_bins x n_bins), sub_bins
        (n_bins x 10) and saved."""
        sc = _capi.WLScheme()
        _capi.check(self._lib.cemc_batch_wl_get_scheme(self.handle, int(chain), ctypes.byref(sc), None, None, None,
                                                       None, None))
        n = int(self.wl_get(chain)["n_bins"])
        edges = np.zeros(n, dtype=np.int32)
        edge_n = np.zeros(n, dtype=np.int32)
        vals = np.zeros((n, n))
        sub = np.zeros((n, 10), dtype=np.uint32)
        saved = np.zeros(n, dtype=np.uint8)
        _capi.check(self._lib.cemc_batch_wl_get_scheme(
            self.handle, int(chain), None, edges.ctypes.data_as(_capi._i32p), edge_n.ctypes.data_as(_capi._i32p),
            vals.ctypes.data_as(_capi._dp), sub.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            saved.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        out = {k: getattr(sc, k) for k, _ in _capi.WLScheme._fields_}
        out.update(edges=edges, edge_n=edge_n, edge_vals=vals, sub_bins=sub, saved=saved)
        return out

    def wl_run(self, n_checks, check_every, log=False, timed=False, shared=False):
        """n_checks rounds of check_every moves on every walker, each followed
        by the flatness test.  shared: every walker visits walker 0's
        histogram and a round's check_every steps are split over the walkers
        (cemc_batch_wl_run_shared).  With log: per-chain accept bits (uint8,
        LSB first); with timed: the device time in ms.  Returns (log, ms)."""
        bits = None
        words = 0
        per = -(-int(check_every) // self.n_chains) if shared else int(check_every)
        if log:
            words = max((int(n_checks) * per + 31) // 32, 1)
            bits = np.zeros((self.n_chains, words), dtype=np.uint32)
        ms = ctypes.c_double(0.0)
        fn = self._lib.cemc_batch_wl_run_shared if shared else self._lib.cemc_batch_wl_run
        _capi.check(fn(
            self.handle, int(n_checks), int(check_every),
            bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if bits is not None else None, words,
            ctypes.byref(ms) if timed else None))
        out_log = bits.view(np.uint8).reshape(self.n_chains, -1) if bits is not None else None
        return out_log, (ms.value if timed else None)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.cemc_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChainPool(ChainBatch):
    """A batch whose chains join and leave while it runs (the resident state
    of a job server).  Capacity is fixed at creation: chain slots, distinct
    models, occupation bytes and canonical tracker entries; ``add`` raises
    ``cemc_amd._capi.PoolFull`` when a chain does not fit.  Runs take one
    move count per slot (0 for slots that should not move)."""

    def __init__(self, max_chains, max_models, occ_bytes, tracker_entries):
        lib = _capi.load()
        self._lib = lib
        h = ctypes.c_void_p()
        _capi.check(lib.cemc_pool_create(int(max_chains), int(max_models), int(occ_bytes), int(tracker_entries),
                                         ctypes.byref(h)))
        self.handle = h
        self.n_chains = int(max_chains)
        self.models = [None] * self.n_chains
        self.n_sites = np.zeros(self.n_chains, dtype=np.int64)
        self.kinds = np.zeros(self.n_chains, dtype=np.int32)
        self.temperatures = np.zeros(self.n_chains)
        self.active = np.zeros(self.n_chains, dtype=bool)

    @staticmethod
    def footprint(model, kind):
        """(occupation bytes, tracker entries) one chain of `model` takes."""
        o, t = ctypes.c_int64(), ctypes.c_int64()
        _capi.check(_capi.load().cemc_pool_footprint(model.handle, int(kind), ctypes.byref(o), ctypes.byref(t)))
        return o.value, t.value

    def add(self, model, chain):
        """Admit one chain (dict as for ChainBatch); returns its slot."""
        d = _capi.ChainDesc()
        keep = []
        _fill_desc(d, model, chain, keep)
        slot = ctypes.c_int32(-1)
        _capi.check(self._lib.cemc_pool_add(self.handle, model.handle, ctypes.byref(d), ctypes.byref(slot)))
        c = slot.value
        self.models[c] = model
        self.n_sites[c] = model.n_sites
        self.kinds[c] = d.kind
        self.temperatures[c] = d.temperature
        self.active[c] = True
        return c

    def remove(self, slot):
        _capi.check(self._lib.cemc_pool_remove(self.handle, int(slot)))
        self.active[slot] = False
        self.models[slot] = None
        self.n_sites[slot] = 0
