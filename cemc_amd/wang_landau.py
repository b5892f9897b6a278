"""Wang-Landau density-of-states sampling on the GPU: the reference's native
sampler (cpp/src/wang_landau_sampler.cpp, driven from
cemc/wanglandau/wang_landau_sampler.py:450-469 ``run_fast_sampler``) for an
atoms object whose calculator is ``cemc_amd.calculator.CE``.

Each walker is one chain of a batch.  A trial move is the reference's
canonical swap, or with ``ensemble="semi-grand-canonical"`` a flip of one
site to another species, whose energy carries the chemical-potential term
E - N sum_d mu_d c1_d (the ensemble of the reference's Python loop,
wang_landau_sampler.py:268-311; the chemical potentials are given per
singlet as for SGCMonteCarlo and folded into the singlet ECIs; for two
species the reference's per-element potentials map onto the singlet's and
the two energies differ by a constant).  The
energy is binned into ``Nbins`` bins over
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

Schemes of the native sampler (``cemc_batch_wl_set_scheme``):

* the inverse-time law (wang_landau_sampler.cpp:344-367): once a flat
  histogram takes more steps than the previous one, f = f_act t_act / t at
  every later flat histogram, with t = steps / Nbins.  The reference times
  convergence in CPU seconds; here it is counted in steps so that runs
  repeat.  ``scheme="inverse_time"`` switches it on for run_fast_sampler
  (the reference's run_fast_sampler leaves it off whatever the scheme; its
  Python-loop InverseTimeScheme is not carried over);
* adaptive energy windows (``mode="adaptive_windows"``,
  adaptive_windows.cpp, wang_landau_sampler.cpp:575-580): the window shrinks
  from the top as its upper part converges, with the walker's lattice
  recorded at the first visit of every bin and sub-bin visit counts kept.

With ``shared_histogram=True`` the walkers visit one histogram, as the
reference's OpenMP walkers do (wang_landau_sampler.cpp:288-298): ln g and
the counts grow by atomic additions in device memory, each check's steps are
split over the walkers, and walker 0 carries the result.  Such runs do not
repeat exactly when f > 0, since a walker reads ln g as the others' visits
land.

Not carried over: the Python ``run`` loop itself (it reads and writes the
reference's database, moves one site at a time from Python and shifts the
energy by the chemical potentials of the two changed elements only) and the
database records.
"""
from __future__ import annotations

import os

import numpy as np

from .batch import CANONICAL, SGC, ChainBatch


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
    :param str scheme: "fixed_f", "square_root_reduction" (both halve f at
        a flat histogram in the native sampler) or "inverse_time"
    :param bool shared_histogram: the walkers visit one histogram (their
        steps per check add up to check_convergence_every)
    :param str ensemble: "canonical" (swaps) or "semi-grand-canonical"
        (flips; the binned energy is E - N sum_d mu_d c1_d, with chem_pot
        given per singlet as for SGCMonteCarlo, e.g. {"c1_0": -0.1})
    """

    def __init__(self, atoms, Nbins=100, initial_f=2.71, flatness_criteria=0.8, fmin=1e-6, Emin=0.0, Emax=1.0,
                 check_convergence_every=1000, n_walkers=1, seed=None, uid=0, scheme="fixed_f",
                 ensemble="canonical", shared_histogram=False, chem_pot=None):
        calc = atoms.get_calculator() if hasattr(atoms, "get_calculator") else getattr(atoms, "calc", None)
        if calc is None or not hasattr(calc, "updater"):
            raise ValueError("the atoms object needs a cemc_amd CE calculator attached")
        if scheme not in ("fixed_f", "inverse_time", "square_root_reduction"):
            raise ValueError("scheme has to be one of ['fixed_f', 'inverse_time', 'square_root_reduction']")
        if ensemble not in ("canonical", "semi-grand-canonical"):
            raise ValueError("Ensemble has to one of ['canonical', 'semi-grand-canonical']")
        sgc = ensemble == "semi-grand-canonical"
        if sgc and not chem_pot:
            raise ValueError("the semi-grand-canonical ensemble needs chemical potentials")
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
        self.scheme = scheme
        self.shared_histogram = bool(shared_histogram)
        self.seed = int.from_bytes(os.urandom(8), "little") if seed is None else int(seed)
        self.uid = int(uid)
        occ = calc.updater.get_occupation()
        self.ensemble = ensemble
        self.chem_pot = dict(chem_pot) if sgc else None
        chains = [dict(kind=SGC if sgc else CANONICAL, temperature=1000.0, occupation=occ, seed=self.seed,
                       uid=self.uid + w, chem_pot=self.chem_pot) for w in range(self.n_walkers)]
        self._batch = ChainBatch([calc.model] * self.n_walkers, chains)
        # the walkers' energy: with the chemical-potential term in SGC
        e0 = float(self._batch.energies()[0]) if sgc else calc.get_energy()
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
        if self.shared_histogram:  # walker 0 holds the group's histogram and f
            self.f_values = [self.walkers[0]["f"]]
            self.converged = bool(self.walkers[0]["converged"])
            self.iter = sum(st["n_moves"] for st in self.walkers)
        else:
            self.f_values = [st["f"] for st in self.walkers]
            self.converged = all(st["converged"] for st in self.walkers)
            self.iter = max(st["n_moves"] for st in self.walkers)
        self.f = max(self.f_values)

    def run_fast_sampler(self, maxsteps=10000000, mode="regular", minimum_window_width=10, sub_bin_file=None,
                         use_inverse_time_algorithm=None):
        """``WangLandauSampler::run(maxsteps)`` on every walker: maxsteps //
        check_convergence_every rounds of check_convergence_every moves, each
        followed by the flatness test (wang_landau_sampler.py:450-469).

        :param mode: "regular" or "adaptive_windows"
        :param minimum_window_width: smallest window of adaptive windows
        :param sub_bin_file: CSV to write walker 0's sub-bin counts to (one
            row of 10 counts per bin; adaptive windows only)
        :param use_inverse_time_algorithm: default: scheme == "inverse_time"
        The schemes start afresh at every call, as the reference builds a
        new sampler per call."""
        if mode not in ("regular", "adaptive_windows"):
            raise ValueError("Unknown mode. Has to one of ['regular', 'adaptive_windows']")
        inv = self.scheme == "inverse_time" if use_inverse_time_algorithm is None else bool(use_inverse_time_algorithm)
        for w in range(self.n_walkers):
            self._batch.wl_set_scheme(w, inverse_time=inv, adaptive=mode == "adaptive_windows",
                                      min_width=int(minimum_window_width))
        if self.shared_histogram and mode == "adaptive_windows":
            raise ValueError("adaptive windows run with one walker per histogram")
        n_outer = int(maxsteps) // self.check_convergence_every
        if n_outer > 0:
            self._batch.wl_run(n_outer, self.check_convergence_every, shared=self.shared_histogram)
        self._pull()
        if sub_bin_file is not None and mode == "adaptive_windows":
            self.save_sub_bin_distribution(sub_bin_file)
        return self.converged

    def scheme_state(self, walker=0):
        """The walker's scheme state (cemc_wl_scheme fields, window cuts,
        sub-bin counts and recorded-state flags)."""
        return self._batch.wl_scheme(walker)

    def sub_bin_distribution(self, walker=0):
        """Visit counts of 10 sub-bins per bin, (Nbins, 10)."""
        return self._batch.wl_scheme(walker)["sub_bins"]

    def save_sub_bin_distribution(self, fname, walker=0):
        """Histogram::save_sub_bin_distribution (histogram.cpp:197-218): one
        line per bin, each count followed by a comma."""
        with open(fname, "w") as out:
            for row in self.sub_bin_distribution(walker):
                out.write("".join("%d," % v for v in row) + "\n")

    def occupation(self, walker=0):
        """Current lattice of one walker (species ids)."""
        return self._batch.occupation(walker)

    def energy(self, walker=0):
        return float(self._batch.energies()[walker])

    def logdos(self, walker=None):
        """ln g per bin: of one walker, or the mean over the walkers of each
        walker's ln g shifted to zero at the lowest bin every walker knows
        (NaN on bins some walker has not visited); with a shared histogram,
        the shared ln g."""
        if self.shared_histogram and walker is None:
            walker = 0
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
