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
