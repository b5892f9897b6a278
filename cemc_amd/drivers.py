"""Drivers over the samplers: simulated annealing in the canonical ensemble
(cemc/mcmc/sa_canonical.py) and a sweep of semi-grand-canonical runs over
temperatures and chemical potentials (cemc/mcmc/mc_parameter_sweep.py).
"""
from __future__ import annotations

import os

import numpy as np

from .mcmc import Montecarlo


class SimulatedAnnealingCanonical:
    """Canonical runs at each temperature in turn, keeping the lowest energy
    met after any move (sa_canonical.py:6-34).  The reference's "maximize"
    mode runs at negative temperatures, where its acceptance rule takes
    every move; the samplers here need positive temperatures, so only
    "minimize" is offered."""

    def __init__(self, atoms, temperatures, mode="minimize", seed=None, uid=0):
        if mode != "minimize":
            raise NotImplementedError("only mode='minimize' runs (the samplers need positive temperatures)")
        self.atoms = atoms
        self.temperatures = list(temperatures)
        self.mode = mode
        self.extremal_energy = 0.0
        self.seed = seed
        self.uid = uid

    def run(self, steps_per_temp=10000):
        """Energies after every move are read back, one launch per move."""
        self.extremal_energy = np.inf
        for k, T in enumerate(self.temperatures):
            seed = None if self.seed is None else int(self.seed) + k
            mc = Montecarlo(self.atoms, T, seed=seed, uid=self.uid)
            mc.update_current_energy()
            trace = mc._energy_trace(int(steps_per_temp))
            mc._sync_back()
            if trace.size:
                self.extremal_energy = min(self.extremal_energy, float(trace.min()))
        return self.extremal_energy


class MCParameterSweep:
    """SGC runs over a list of {"temperature", "chemical_potential"}
    parameter sets, each after ``reset`` and with equilibration, collecting
    ``get_thermodynamic`` (mc_parameter_sweep.py:6-103).  The reference
    appends the results to an HDF5 file through h5py; here they go to an
    ``.npz`` file (one array per quantity, appended to an existing file)."""

    known_parameters = {"MonteCarlo": ["temperature", "composition"],
                        "SGCMonteCarlo": ["temperature", "chemical_potential"]}

    def __init__(self, parameters, mc_obj, nsteps=100000, data_getter=None, outfile="default_output.npz",
                 equil_params=None):
        self.parameters = parameters
        self.mc_obj = mc_obj
        self.nsteps = int(nsteps)
        self.data_getter = data_getter
        self.outfile = outfile
        self.equil_params = equil_params
        self.check_initialization()

    def check_initialization(self):
        """mc_parameter_sweep.py:21-51."""
        if self.mc_obj.name not in self.known_parameters:
            raise ValueError("Monte Carlo instance was not recognized. Known MonteCarlo object: {}"
                             .format(list(self.known_parameters)))
        for i, p in enumerate(self.parameters):
            for req in self.known_parameters[self.mc_obj.name]:
                if req not in p:
                    raise ValueError("Required parameter {} not given for entry {}.".format(req, i))
            if "chemical_potential" in p and not isinstance(p["chemical_potential"], dict):
                raise ValueError("Chemical potential has to be given as a dictionary")
            if "temperature" in p:
                try:
                    float(p["temperature"])
                except (TypeError, ValueError):
                    raise ValueError("Temperature has to be given as a float") from None

    def run(self):
        """mc_parameter_sweep.py:53-77."""
        if self.mc_obj.name != "SGCMonteCarlo":
            raise NotImplementedError("Parameter sweep for the MC object not supported yet!")
        all_data = []
        for p in self.parameters:
            self.mc_obj.reset()
            self.mc_obj.T = float(p["temperature"])
            self.mc_obj.chemical_potential = p["chemical_potential"]
            self.mc_obj.runMC(steps=self.nsteps, equil=True, equil_params=self.equil_params)
            data = self.mc_obj.get_thermodynamic()
            if self.data_getter is not None:
                data.update(dict(self.data_getter(self.mc_obj)))
            all_data.append(data)
        if self.outfile:
            self.save(all_data)
        return all_data

    def save(self, data):
        """One array per quantity, appended to what the file holds."""
        ignore = {"timestamp", "python_version"}
        flat = {k: np.array([d[k] for d in data]) for k in data[0] if k not in ignore}
        if os.path.exists(self.outfile):
            with np.load(self.outfile) as old:
                for k in old.files:
                    if k in flat:
                        flat[k] = np.concatenate([old[k], flat[k]])
        np.savez(self.outfile, **flat)
