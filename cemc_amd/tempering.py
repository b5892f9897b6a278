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


class ParallelTempering:
    """Replicas of a ``cemc_amd.mcmc.Montecarlo`` object over a temperature
    scheme (parallel_tempering.py:5-20).

    The scheme is ``temperatures`` when given, else the first column of
    ``temp_scheme_file`` when it exists (parallel_tempering.py:32-45), else
    ``n_replicas`` temperatures log-spaced from Tmax down to Tmin.  (The
    reference's adaptive scheme search, :47-136, is not carried over.)
    """

    def __init__(self, mc_obj=None, Tmax=1500.0, Tmin=100.0, temp_scheme_file="temp_scheme.csv",
                 temperatures=None, n_replicas=16, seed=None, uid=None):
        from .mcmc import Montecarlo
        if not isinstance(mc_obj, Montecarlo):
            raise TypeError("mc_obj has to be of type Montecarlo!")
        if temperatures is None and temp_scheme_file and os.path.exists(temp_scheme_file):
            data = np.loadtxt(temp_scheme_file, delimiter=",", ndmin=2)
            temperatures = data[:, 0].tolist()
        if temperatures is None:
            temperatures = np.geomspace(Tmax, Tmin, int(n_replicas)).tolist()
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
        (parallel_tempering.py:177-191).  Returns the device time in ms."""
        mc_args = dict(mc_args or {})
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
