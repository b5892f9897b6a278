"""Parallel tempering on the GPU: the reference's ``ParallelTempering``
(cemc/mcmc/parallel_tempering.py:5-191) with every replica resident on the
device and the exchange step done by a kernel after each cycle's sweep.

Replica slot t has temperature T_t; exchanges swap which chain (lattice)
holds which slot, so configurations never move.  A cycle is one
``runMC(steps)`` per replica followed by one exchange step over even or odd
neighbour pairs (direction drawn per cycle); the stream definitions are in
include/cemc_gpu.h.  Sweep and exchange launches of up to 16 cycles are
captured in a HIP graph and replayed.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi
from .batch import CANONICAL, ChainBatch


class ReplicaExchange:
    """Exchange state over all chains of a ChainBatch (slot t starts with chain t)."""

    def __init__(self, batch, seed=0, uid=0):
        self.batch = batch
        self._lib = _capi.load()
        h = ctypes.c_void_p()
        _capi.check(self._lib.cemc_pt_create(batch.handle, int(seed) & 0xFFFFFFFFFFFFFFFF,
                                             int(uid) & 0xFFFFFFFFFFFFFFFF, ctypes.byref(h)))
        self.handle = h
        self.n_rep = batch.n_chains

    def run(self, n_cycles, moves_per_cycle, use_graph=True):
        """Run cycles; returns the device time in ms."""
        ms = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_pt_run(self.handle, int(n_cycles), int(moves_per_cycle), int(bool(use_graph)),
                                          ctypes.byref(ms)))
        return ms.value

    def state(self):
        R = self.n_rep
        cos = np.zeros(R, dtype=np.int32)
        temps = np.zeros(R)
        acc = np.zeros(max(R - 1, 1), dtype=np.int64)
        tries = np.zeros(max(R - 1, 1), dtype=np.int64)
        n = np.zeros(R, dtype=np.int64)
        mean = np.zeros(R)
        var = np.zeros(R)
        _capi.check(self._lib.cemc_pt_get(self.handle, cos.ctypes.data_as(_capi._i32p), temps.ctypes.data_as(_capi._dp),
                                          acc.ctypes.data_as(_capi._i64p), tries.ctypes.data_as(_capi._i64p),
                                          n.ctypes.data_as(_capi._i64p), mean.ctypes.data_as(_capi._dp),
                                          var.ctypes.data_as(_capi._dp)))
        return {"chain_of_slot": cos, "temperatures": temps, "n_accept": acc[:R - 1], "n_try": tries[:R - 1],
                "n_samples": n, "mean_energy": mean, "var_energy": var}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.cemc_pt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _accept_probability(E1, E2, T1, T2):
    """parallel_tempering.py:138-144 (not capped at 1)."""
    from .lattice import KB_EV
    return float(np.exp((1.0 / (KB_EV * T1) - 1.0 / (KB_EV * T2)) * (E1 - E2)))


def find_temperature_scheme(mc_obj, Tmax, Tmin, target_accept=0.2, log=None):
    """Temperature scheme of the reference's search
    (parallel_tempering.py:47-136), each trial a GPU launch of 10 N moves on
    a chain of its own.  The first replica runs 10 N moves at Tmax; each next
    temperature starts from a copy of the previous replica's lattice: the
    trial temperature is halved until the estimated exchange acceptance
    exp((1/kT_prev - 1/kT) (E_prev - E)) drops to target_accept, then
    bisected until the acceptance is within 0.01 of the target or the
    interval is below 1e-4 K.  Every trial is one runMC(steps=10 N,
    equil=False), which resets the averages before it samples
    (montecarlo.py:796-800): E is the mean energy of that trial's moves only,
    E_prev the mean of the previous replica's last trial.  (The reference's
    runMC also makes one move before and 1000 energy-bias probe moves,
    montecarlo.py:762, 797-799, before the sampled moves; they are not made
    here.)  Stops at Tmin, or when halving reaches Tmin without a candidate.
    Returns (temperatures, acceptances), the first acceptance 0."""
    log = log or (lambda msg: None)
    natoms = len(mc_obj.atoms)
    nsteps = 10 * natoms
    model = mc_obj.calc.model
    extra = mc_obj._chain_extra()
    uid0 = int(mc_obj.uid)

    def chain(T, occ, r):
        d = dict(kind=mc_obj.kind, temperature=float(T), occupation=occ, seed=mc_obj.seed, uid=uid0 + r)
        d.update(extra)
        return ChainBatch([model], [d])

    def run_fixed(ch, T):
        # runMC(steps=nsteps, equil=False): averages over this run's moves only
        ch.set_temperature(0, T)
        ch.reset_stats(0)
        ch.run(nsteps)
        return ch.thermo(0)["mean_energy"]

    cur = chain(Tmax, mc_obj.calc.updater.get_occupation(), 0)
    cur_E = run_fixed(cur, Tmax)
    T_cur = float(Tmax)
    temps, accepts = [T_cur], [0.0]
    lowest = T_cur
    r = 1
    while lowest > Tmin:
        new = chain(T_cur, cur.occupation(0), r)
        trial = T_cur / 2.0
        accept = 1.0
        found = False
        while accept > target_accept and trial > Tmin:
            accept = _accept_probability(cur_E, run_fixed(new, trial), T_cur, trial)
            found = accept <= target_accept
            trial /= 2.0
        if not found:
            new.close()
            break
        upper, lower = T_cur, trial
        while True:
            new_T = 0.5 * (upper + lower)
            new_E = run_fixed(new, new_T)
            new_accept = _accept_probability(cur_E, new_E, T_cur, new_T)
            if new_accept > target_accept:
                upper = new_T
            else:
                lower = new_T
            if abs(new_accept - target_accept) < 0.01 or upper - lower < 1e-4:
                break
        log("temperature %d: %.4f K (exchange acceptance %.3f)" % (r, new_T, new_accept))
        temps.append(new_T)
        accepts.append(new_accept)
        cur.close()
        cur, T_cur, lowest, cur_E = new, new_T, new_T, new_E
        r += 1
    cur.close()
    return temps, accepts


class ParallelTempering:
    """Replicas of a ``cemc_amd.mcmc.Montecarlo`` object over a temperature
    scheme (parallel_tempering.py:5-20).

    The scheme is ``temperatures`` when given, else the first column of
    ``temp_scheme_file`` when it exists (parallel_tempering.py:32-45), else
    ``n_replicas`` temperatures log-spaced from Tmax down to Tmin when
    ``n_replicas`` is given, else the reference's search
    (``find_temperature_scheme``, parallel_tempering.py:47-136), whose
    result is written to ``temp_scheme_file`` as the reference does.
    """

    def __init__(self, mc_obj=None, Tmax=1500.0, Tmin=100.0, temp_scheme_file="temp_scheme.csv",
                 temperatures=None, n_replicas=None, seed=None, uid=None, target_accept=0.2):
        from .mcmc import Montecarlo
        if not isinstance(mc_obj, Montecarlo):
            raise TypeError("mc_obj has to be of type Montecarlo!")
        if temperatures is None and temp_scheme_file and os.path.exists(temp_scheme_file):
            data = np.loadtxt(temp_scheme_file, delimiter=",", ndmin=2)
            temperatures = data[:, 0].tolist()
        if temperatures is None and n_replicas is not None:
            temperatures = np.geomspace(Tmax, Tmin, int(n_replicas)).tolist()
        if temperatures is None:
            temperatures, accepts = find_temperature_scheme(mc_obj, Tmax, Tmin, target_accept=target_accept)
            if temp_scheme_file:
                np.savetxt(temp_scheme_file, np.vstack((temperatures, accepts)).T, delimiter=",",
                           header="Temperature (K), Acceptance probabability")
        temperatures = [float(t) for t in temperatures]
        if len(temperatures) < 2:
            raise ValueError("parallel tempering needs at least two temperatures")
        self.mc = mc_obj
        self.Tmax, self.Tmin = max(temperatures), min(temperatures)
        self.natoms = len(mc_obj.atoms)
        occ = mc_obj.calc.updater.get_occupation()
        seed = mc_obj.seed if seed is None else int(seed)
        uid = mc_obj.uid if uid is None else int(uid)
        model = mc_obj.calc.model
        chains = []
        for r, T in enumerate(temperatures):
            d = dict(kind=mc_obj.kind, temperature=T, occupation=occ, seed=seed, uid=uid + r)
            d.update(mc_obj._chain_extra())
            chains.append(d)
        self.batch = ChainBatch([model] * len(chains), chains)
        self.exchange = ReplicaExchange(self.batch, seed=seed, uid=(uid + len(chains)) | (1 << 63))
        self._temps = temperatures

    @property
    def temperature_scheme(self):
        return list(self._temps)

    def run(self, mc_args=None, num_exchange_cycles=10, use_graph=True):
        """num_exchange_cycles cycles of ``mc_args["steps"]`` trial moves per
        replica (default: one sweep) then one exchange step
        (parallel_tempering.py:177-191).  Returns the device time in ms.

        The reference passes ``mc_args`` to a full ``runMC`` per replica and
        cycle; here a cycle is exactly ``steps`` moves per replica inside a
        captured graph: no per-cycle equilibration (the reference's default
        ``equil=True``), energy-bias probe or precision mode.  ``equil`` may
        be given as False and ``mode`` as "fixed"; any other key, or
        ``equil=True`` given explicitly, raises NotImplementedError rather
        than being ignored."""
        mc_args = dict(mc_args or {})
        unsupported = sorted(k for k in mc_args if k not in ("steps", "equil", "mode", "verbose"))
        if unsupported:
            raise NotImplementedError("ParallelTempering.run supports mc_args steps, equil=False and "
                                      "mode='fixed' only; got %s" % ", ".join(unsupported))
        if mc_args.get("equil", False):
            raise NotImplementedError("per-cycle equilibration (equil=True) is not run on the device; "
                                      "pass equil=False")
        if mc_args.get("mode", "fixed") != "fixed":
            raise NotImplementedError("only mode='fixed' cycles run on the device")
        steps = int(mc_args.get("steps", self.natoms))
        return self.exchange.run(num_exchange_cycles, steps, use_graph=use_graph)

    def get_thermodynamic(self):
        """Per temperature: mean energy, heat capacity and exchange acceptance
        with the next temperature (pairs in scheme order)."""
        from .lattice import KB_EV
        st = self.exchange.state()
        out = []
        for t, T in enumerate(self._temps):
            rate = None
            if t < len(self._temps) - 1 and st["n_try"][t] > 0:
                rate = float(st["n_accept"][t]) / float(st["n_try"][t])
            out.append({"temperature": T, "energy": float(st["mean_energy"][t]),
                        "heat_capacity": float(st["var_energy"][t]) / (KB_EV * T * T),
                        "n_samples": int(st["n_samples"][t]), "exchange_accept_rate": rate,
                        "chain": int(st["chain_of_slot"][t])})
        return out

    def replica_occupation(self, slot):
        """Occupation (species ids) of the lattice at temperature slot `slot`."""
        st = self.exchange.state()
        return self.batch.occupation(int(st["chain_of_slot"][slot]))
