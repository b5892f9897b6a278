"""Wang-Landau density-of-states sampling on the GPU: the reference's native
sampler (cpp/src/wang_landau_sampler.cpp, driven from
cemc/wanglandau/wang_landau_sampler.py:450-469 ``run_fast_sampler``) for an
atoms object whose calculator is ``cemc_amd.calculator.CE``.

Each walker is one canonical chain of a batch.  A trial move is the
reference's canonical swap; its energy is binned into ``Nbins`` bins over
[Emin, Emax) and the swap is kept with probability
min(1, g(current) / g(proposed)), where ln g is raised by f at every visit
(``WangLandauSampler::step``, :232-317).  After every
``check_convergence_every`` moves the histogram of visits is tested for
flatness (min > flatness_criteria * mean over the bins visited so far,
``Histogram::is_flat``); a flat histogram halves f and clears the counts,
and the walker stops once f < fmin (``run``, :319-404).  Random numbers come
from the Philox4x64-10 stream keyed by (seed, uid + walker).

The reference runs its OpenMP walkers on one shared histogram under a lock.
Here every walker keeps its own histogram, so the walkers are independent
fixed-f runs of the same system (the reference's fixed_f scheme asks for
several runs whose ln g are averaged: wang_landau_sampler.py:55-57); they
run in one launch, one wavefront each, with ln g and the counts in LDS.

Not carried over: the inverse-time scheme, adaptive energy windows
(adaptive_windows.cpp), sub-bin distributions, the Python ``run`` loop with
its SGC moves, and the database records.
"""
from __future__ import annotations

import os

import numpy as np

from .batch import CANONICAL, ChainBatch


class Histogram:
    """Python-side view of one walker's histogram (the attributes of
    cemc/wanglandau/histogram.py:12-24 that the native sampler reads and
    writes back)."""

    def __init__(self, Nbins, Emin, Emax):
        self.Nbins = int(Nbins)
        self.Emin = float(Emin)
        self.Emax = float(Emax)
        self.histogram = np.zeros(self.Nbins, dtype=np.int32)
        self.logdos = np.zeros(self.Nbins)
        self.known_state = np.zeros(self.Nbins, dtype=np.uint8)

    def get_energy(self, indx):
        """Lower edge of bin indx (histogram.py:30-34)."""
        return self.Emin + (self.Emax - self.Emin) * indx / self.Nbins

    def get_bin(self, energy):
        """histogram.cpp:66-69: (E - Emin) * Nbins / (Emax - Emin), truncated."""
        return int((energy - self.Emin) * self.Nbins / (self.Emax - self.Emin))


class WangLandau:
    """Wang-Landau walkers on copies of ``atoms``.

    :param atoms: atoms object with a ``cemc_amd.calculator.CE`` attached
    :param int Nbins: bins of the histogram
    :param float initial_f: modification factor f at the start
    :param float flatness_criteria: flat when min > criteria * mean
    :param float fmin: the walkers stop once f < fmin
    :param float Emin, Emax: energy range of the histogram (total energy, eV)
    :param int check_convergence_every: moves between flatness tests
    :param int n_walkers: independent walkers (chains), all started from atoms
    :param seed: Philox key word 0 (default: drawn from the OS)
    :param int uid: Philox key word 1 of walker 0 (walker w uses uid + w)
    """

    def __init__(self, atoms, Nbins=100, initial_f=2.71, flatness_criteria=0.8, fmin=1e-6, Emin=0.0, Emax=1.0,
                 check_convergence_every=1000, n_walkers=1, seed=None, uid=0, scheme="fixed_f",
                 ensemble="canonical"):
        calc = atoms.get_calculator() if hasattr(atoms, "get_calculator") else getattr(atoms, "calc", None)
        if calc is None or not hasattr(calc, "updater"):
            raise ValueError("the atoms object needs a cemc_amd CE calculator attached")
        if scheme != "fixed_f":
            raise NotImplementedError("only the fixed_f scheme runs on the GPU")
        if ensemble != "canonical":
            raise NotImplementedError("the GPU walkers make canonical swaps")
        if not Emax > Emin:
            raise ValueError("Emax must exceed Emin")
        if int(Nbins) < 1 or int(n_walkers) < 1 or int(check_convergence_every) < 1:
            raise ValueError("Nbins, n_walkers and check_convergence_every must be positive")
        self.atoms = atoms
        self.calc = calc
        self.Nbins = int(Nbins)
        self.Emin = float(Emin)
        self.Emax = float(Emax)
        self.f = float(initial_f)
        self.fmin = float(fmin)
        self.flatness_criteria = float(flatness_criteria)
        self.check_convergence_every = int(check_convergence_every)
        self.n_walkers = int(n_walkers)
        self.seed = int.from_bytes(os.urandom(8), "little") if seed is None else int(seed)
        self.uid = int(uid)
        occ = calc.updater.get_occupation()
        chains = [dict(kind=CANONICAL, temperature=1000.0, occupation=occ, seed=self.seed, uid=self.uid + w)
                  for w in range(self.n_walkers)]
        self._batch = ChainBatch([calc.model] * self.n_walkers, chains)
        e0 = calc.get_energy()
        b0 = Histogram(self.Nbins, self.Emin, self.Emax).get_bin(e0)
        self.current_bin = [b0 if 0 <= b0 < self.Nbins else 0] * self.n_walkers
        for w in range(self.n_walkers):
            self._batch.wl_set(w, self.Nbins, self.Emin, self.Emax, self.f, self.fmin, self.flatness_criteria,
                               current_bin=self.current_bin[w], is_first=True)
        self.histograms = [Histogram(self.Nbins, self.Emin, self.Emax) for _ in range(self.n_walkers)]
        self.converged = False
        self.iter = 0
        self._pull()

    @property
    def histogram(self):
        """Histogram of walker 0."""
        return self.histograms[0]

    def _pull(self):
        self.walkers = []
        for w in range(self.n_walkers):
            st = self._batch.wl_get(w)
            h = self.histograms[w]
            h.histogram = st["hist"]
            h.logdos = st["logdos"]
            h.known_state = st["known"]
            self.current_bin[w] = st["current_bin"]
            self.walkers.append(st)
        self.f_values = [st["f"] for st in self.walkers]
        self.f = max(self.f_values)
        self.converged = all(st["converged"] for st in self.walkers)
        self.iter = max(st["n_moves"] for st in self.walkers)

    def run_fast_sampler(self, maxsteps=10000000, mode="regular"):
        """``WangLandauSampler::run(maxsteps)`` on every walker: maxsteps //
        check_convergence_every rounds of check_convergence_every moves, each
        followed by the flatness test (wang_landau_sampler.py:450-469)."""
        if mode != "regular":
            raise NotImplementedError("adaptive energy windows are not available on the GPU path")
        n_outer = int(maxsteps) // self.check_convergence_every
        if n_outer > 0:
            self._batch.wl_run(n_outer, self.check_convergence_every)
        self._pull()
        return self.converged

    def occupation(self, walker=0):
        """Current lattice of one walker (species ids)."""
        return self._batch.occupation(walker)

    def energy(self, walker=0):
        return float(self._batch.energies()[walker])

    def logdos(self, walker=None):
        """ln g per bin: of one walker, or the mean over the walkers of each
        walker's ln g shifted to zero at the lowest bin every walker knows
        (NaN on bins some walker has not visited)."""
        if walker is not None:
            h = self.histograms[walker]
            out = np.where(h.known_state != 0, h.logdos, np.nan)
            return out
        known = np.all([h.known_state != 0 for h in self.histograms], axis=0)
        if not known.any():
            return np.full(self.Nbins, np.nan)
        ref = int(np.argmax(known))
        rows = [h.logdos - h.logdos[ref] for h in self.histograms]
        return np.where(known, np.mean(rows, axis=0), np.nan)

    def bin_energies(self):
        """Centre of every bin."""
        w = (self.Emax - self.Emin) / self.Nbins
        return self.Emin + w * (np.arange(self.Nbins) + 0.5)

    def close(self):
        if self._batch is not None:
            self._batch.close()
            self._batch = None
