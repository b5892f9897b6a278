"""Linear vibrational correction to the ECIs (cemc/mcmc/linear_vib_correction.py):
ECI_k(T) = ECI_k + v_k kB T for the clusters k in ``eci_per_kbT``.

The calculator adds the terms to its ECIs in place at a sampler's
temperature (``CE.include_linvib_in_ecis``), and the samplers report energies
without them (``Montecarlo.current_energy_without_vib``,
montecarlo.py:452-459).  The Boltzmann constant is the one the reference's
updater uses for the vibrational energy (cpp/src/linear_vib_correction.cpp:3).
"""
from __future__ import annotations

import copy

KB_VIB = 8.6173303e-5


class LinearVibCorrection:
    """Same state and methods as the reference class
    (linear_vib_correction.py:6-76): ``include`` adds v_k kB T to a dict of
    ECIs in place once (a second call only records the temperature and
    returns None), ``reset`` removes them at the last recorded temperature."""

    def __init__(self, eci_per_kbT):
        self.eci_per_kbT = dict(eci_per_kbT)
        self.orig_eci = None
        self.current_eci = None
        self.vibs_included = False
        self.temperature = 0.0

    def check_provided_eci_match(self, provided_eci):
        """Kept for the interface; the reference returns before checking
        (linear_vib_correction.py:14-34)."""
        return None

    def include(self, eci_with_out_vib, T):
        """linear_vib_correction.py:36-51."""
        self.temperature = T
        if self.vibs_included:
            return None
        for key in self.eci_per_kbT:
            if key not in eci_with_out_vib:
                raise KeyError("The cluster {} is not in the original ECIs!".format(key))
        for key, value in self.eci_per_kbT.items():
            eci_with_out_vib[key] += value * KB_VIB * T
        self.vibs_included = True
        self.current_eci = copy.deepcopy(eci_with_out_vib)
        return eci_with_out_vib

    def reset(self, eci_with_vibs):
        """linear_vib_correction.py:53-62."""
        if not self.vibs_included:
            return None
        for key, value in self.eci_per_kbT.items():
            eci_with_vibs[key] -= value * KB_VIB * self.temperature
        self.current_eci = copy.deepcopy(eci_with_vibs)
        self.vibs_included = False
        return eci_with_vibs

    def energy(self, T, cf):
        """Vibrational energy per site, sum_k v_k cf_k kB T in the name
        order of the reference updater's map (linear_vib_correction.cpp:8-16).
        The reference's Python ``energy_due_to_vibrations`` iterates the
        dict without ``items()`` and cannot run; this is the updater's form."""
        e = 0.0
        for key in sorted(self.eci_per_kbT):
            e += cf[key] * self.eci_per_kbT[key] * KB_VIB * T
        return e

    energy_due_to_vibrations = energy
