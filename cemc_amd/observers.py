"""Observers for ``Montecarlo.attach`` (cemc/mcmc/mc_observers.py): callables
that receive the lattice changes since their previous call, with the atoms,
the calculator and the sampler's ``current_energy`` already at the state
after those changes.  ``get_thermodynamic`` adds each attached observer's
``get_averages()`` to its result, as the reference does
(montecarlo.py:885-887).

The SGCObserver's sums (mc_observers.py:185-290) are kept on the device by
every chain and are not an observer here.
"""
from __future__ import annotations

import copy

import numpy as np


class MCObserver:
    """Base class (mc_observers.py:12-37)."""

    def __init__(self):
        self.name = "GenericObserver"

    def __call__(self, system_changes):
        pass

    def reset(self):
        pass

    def get_averages(self):
        return {}


class CorrelationFunctionTracker(MCObserver):
    """History of the correlation functions (mc_observers.py:40-78)."""

    def __init__(self, ce_calc):
        self.cf = []
        self.ce_calc = ce_calc
        self.name = "CorrelationFunctionTracker"

    def __call__(self, system_changes):
        self.cf.append(copy.deepcopy(self.ce_calc.cf))


class PairCorrelationObserver(MCObserver):
    """Averages and standard deviations of the pair correlation functions
    over the calls (mc_observers.py:81-135)."""

    def __init__(self, ce_calc):
        self.ce_calc = ce_calc
        if ce_calc.updater is None:
            raise RuntimeError("This observer can only be used with the C++ version of the CF updater")
        self.n_entries = 0
        self.name = "PairCorrelationObserver"
        self.cf = {k: 0.0 for k in ce_calc.eci if k.startswith("c2_")}
        self.cf_squared = {k: 0.0 for k in self.cf}

    def __call__(self, system_changes):
        new_cf = self.ce_calc.updater.get_cf()
        self.n_entries += 1
        for key in self.cf:
            self.cf[key] += new_cf[key]
            self.cf_squared[key] += new_cf[key] ** 2

    def get_averages(self):
        return {key: value / self.n_entries for key, value in self.cf.items()}

    def get_std(self):
        return {key: float(np.sqrt(self.cf_squared[key] / self.n_entries - (self.cf[key] / self.n_entries) ** 2))
                for key in self.cf}


class LowestEnergyStructure(MCObserver):
    """Lowest-energy state seen at the calls (mc_observers.py:138-182)."""

    def __init__(self, ce_calc, mc_obj, verbose=False):
        self.ce_calc = ce_calc
        self.mc_obj = mc_obj
        self.lowest_energy = np.inf
        self.lowest_energy_cf = None
        self.atoms = None
        self.lowest_energy_atoms = None
        self.name = "LowestEnergyStructure"
        self.verbose = verbose

    def __call__(self, system_changes):
        e = self.mc_obj.current_energy
        if self.atoms is None or self.lowest_energy_cf is None or e < self.lowest_energy:
            d_e = e - self.lowest_energy
            first = self.atoms is None
            self.lowest_energy = e
            self.lowest_energy_cf = self.ce_calc.get_cf()
            self.atoms = self.mc_obj.atoms.copy()
            self.lowest_energy_atoms = self.atoms
            if self.verbose and not first:
                print("Found new low energy structure. New energy: {} eV. Change: {} eV".format(e, d_e))


class EnergyEvolution(MCObserver):
    """Energies (without the vibrational part) at the calls
    (mc_observers.py:688-706)."""

    def __init__(self, mc_obj):
        MCObserver.__init__(self)
        self.mc = mc_obj
        self.energies = []
        self.name = "EnergyEvolution"

    def __call__(self, system_changes):
        self.energies.append(self.mc.current_energy_without_vib())

    def reset(self):
        self.energies = []

    def save(self, fname="energy_evolution.csv"):
        np.savetxt(fname, self.energies, delimiter=",")


class SiteOrderParameter(MCObserver):
    """Number of sites whose symbol differs from the symbols at creation,
    averaged over the calls (mc_observers.py:614-685); symbols stand in for
    the atomic numbers the reference compares."""

    def __init__(self, atoms):
        self.atoms = atoms
        self.orig_symbols = np.array([a.symbol for a in atoms])
        self.avg_num_changed = 0
        self.avg_num_changed_sq = 0
        self.num_calls = 0
        self.current_num_changed = 0
        self.site_changed = np.zeros(len(atoms), dtype=bool)

    def _check_all_sites(self):
        now = np.array([a.symbol for a in self.atoms])
        self.site_changed = now != self.orig_symbols
        self.current_num_changed = int(np.count_nonzero(self.site_changed))

    def reset(self):
        self.avg_num_changed = 0
        self.avg_num_changed_sq = 0
        self.num_calls = 0
        self.current_num_changed = 0
        self._check_all_sites()

    def __call__(self, system_changes):
        self.num_calls += 1
        for change in system_changes:
            i = int(change[0])
            changed = self.atoms[i].symbol != self.orig_symbols[i]
            if changed != self.site_changed[i]:
                self.current_num_changed += 1 if changed else -1
                self.site_changed[i] = changed
        self.avg_num_changed += self.current_num_changed
        self.avg_num_changed_sq += self.current_num_changed ** 2

    def get_averages(self):
        average = float(self.avg_num_changed) / self.num_calls
        var = max(float(self.avg_num_changed_sq) / self.num_calls - average ** 2, 0.0)
        return {"site_order_average": average, "site_order_std": float(np.sqrt(var))}


class EnergyHistogram(MCObserver):
    """Histogram of the energies (without the vibrational part) at the
    calls: the first ``buffer_size`` energies fix the range, later ones are
    binned (mc_observers.py:709-760; the reference sizes the histogram with
    ``len(n_bins)``, which cannot run, where ``n_bins`` is meant).  Later
    energies outside that range, which the reference's indexing would wrap
    or fail on, are counted in ``n_outside`` instead."""

    def __init__(self, mc_obj, buffer_size=100000, n_bins=100):
        MCObserver.__init__(self)
        self.mc = mc_obj
        self.buffer = np.zeros(buffer_size)
        self.n_bins = int(n_bins)
        self._next = 0
        self._histogram = None
        self.Emin = None
        self.Emax = None
        self.sample_in_buffer = True
        self.n_outside = 0

    def __call__(self, system_changes):
        e = self.mc.current_energy_without_vib()
        if self.sample_in_buffer:
            self.buffer[self._next] = e
            self._next += 1
            if self._next >= len(self.buffer):
                self._on_buffer_full()
        elif self.Emin <= e <= self.Emax:
            self._histogram[self._get_indx(e)] += 1
        else:
            self.n_outside += 1

    def _on_buffer_full(self):
        sample = self.buffer[:self._next]
        self.Emin = float(np.min(sample))
        self.Emax = float(np.max(sample))
        self._histogram = np.zeros(self.n_bins)
        if self.Emax == self.Emin:
            self._histogram[0] = len(sample)
            self.sample_in_buffer = False
            return
        for e in sample:
            self._histogram[self._get_indx(e)] += 1
        self.sample_in_buffer = False

    def _get_indx(self, e):
        """int((E - Emin)(n_bins - 1)/(Emax - Emin)), as the reference bins."""
        if self.Emin is None or self.Emax is None:
            raise RuntimeError("This function should never be called before the histogram has been updated "
                               "at least once!")
        return int((e - self.Emin) * (self.n_bins - 1) / (self.Emax - self.Emin))

    @property
    def histogram(self):
        if self._histogram is None:
            self._on_buffer_full()
        return self._histogram


class MCBackup(MCObserver):
    """Write the sampler with ``save`` at every call (mc_observers.py:763-805).
    The reference's optional database table needs the ``dataset`` package;
    a ``db_name`` raises NotImplementedError."""

    def __init__(self, mc_obj, backup_file="montecarlo_backup.json", db_name="", db_tab_name="mc_backup",
                 db_id=None, overwrite_db_row=True):
        MCObserver.__init__(self)
        if db_name:
            raise NotImplementedError("database backups need the dataset package")
        self.mc_obj = mc_obj
        self.backup_file = backup_file
        self.name = "MCBackup"

    def __call__(self, system_changes):
        self.mc_obj.save(self.backup_file)
