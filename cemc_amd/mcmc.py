"""Metropolis samplers on the GPU: the reference's ``Montecarlo``
(cemc/mcmc/montecarlo.py:39-1038) and ``SGCMonteCarlo``
(cemc/mcmc/sgc_montecarlo.py:12-448) for an atoms object whose calculator is
``cemc_amd.calculator.CE``.

``runMC(steps=...)`` advances the chain on the GPU in one launch: each trial
move is the reference's ``_mc_step`` (proposal of ``_get_trial_move``,
Metropolis test of ``_accept``, commit or undo, species-tracker update) with
its random numbers taken from a Philox4x64-10 stream keyed by
(seed, uid); move t uses counter (t, 0, 0, 0), and the counter carries on
across calls.  Energy sums (and, in the SGC ensemble, the SGCObserver's
singlet sums, cemc/mcmc/mc_observers.py:222-270) are kept per move on the
device; ``get_thermodynamic`` turns them into the reference's quantities.

Not carried over (listed as later work): the adaptive equilibration and
correlation-time estimates (montecarlo.py:461-697), the ``"prec"`` stopping
mode, waste recycling, constraints, bias potentials and Python observers
called per move.  ``equil=True`` runs a fixed burn-in instead (see runMC).
"""
from __future__ import annotations

import numpy as np

from .batch import CANONICAL, SGC, ChainBatch
from .lattice import KB_EV


def _fresh_seed():
    return int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])


class Montecarlo:
    """Canonical (fixed composition) Metropolis sampling with swap moves.

    :param atoms: atoms object with a ``cemc_amd.calculator.CE`` attached
    :param float temp: temperature in kelvin
    :param seed: Philox key word 0 (default: drawn from the OS)
    :param int uid: Philox key word 1 (default 0)
    """

    kind = CANONICAL

    def __init__(self, atoms, temp, indeces=None, logfile="", plot_debug=False, min_acc_rate=0.0,
                 recycle_waste=False, max_constraint_attempts=10000, accept_first_trial_move_after_reset=False,
                 seed=None, uid=0):
        calc = atoms.get_calculator() if hasattr(atoms, "get_calculator") else getattr(atoms, "calc", None)
        if calc is None or not hasattr(calc, "updater"):
            raise ValueError("the atoms object needs a cemc_amd CE calculator attached")
        if recycle_waste:
            raise NotImplementedError("waste recycling is not available on the GPU path")
        if max_constraint_attempts <= 0:
            raise ValueError("Max. constraint attempts has to be > 0!")
        self.name = "MonteCarlo"
        self.atoms = atoms
        self.calc = calc
        self._T = float(temp)
        self.min_acc_rate = min_acc_rate
        self.indeces = range(len(atoms)) if indeces is None else indeces
        self.seed = _fresh_seed() if seed is None else int(seed)
        self.uid = int(uid)
        self.symbols = sorted(set(a.symbol for a in atoms))
        self.current_step = 0
        self.num_accepted = 0
        self._moves = 0          # Philox counter of the next move
        self._batch = None
        self._batch_occ = None   # occupation the batch was built from / last synced
        self._batch_eci = None
        self._chem_pot_vec = None
        self.current_energy = calc.get_energy()

    # ------------------------------------------------------------ properties
    @property
    def T(self):
        return self._T

    @T.setter
    def T(self, value):
        self._T = float(value)
        if self._batch is not None:
            self._batch.set_temperature(0, self._T)

    # -------------------------------------------------------------- helpers
    def _chain_extra(self):
        return {}

    def _ensure_batch(self):
        occ = self.calc.updater.get_occupation()
        eci = tuple(sorted(self.calc.eci.items()))
        if self._batch is not None and np.array_equal(occ, self._batch_occ) and eci == self._batch_eci:
            return self._batch
        if self._batch is not None:
            self._batch.close()
        chain = dict(kind=self.kind, temperature=self._T, occupation=occ, seed=self.seed, uid=self.uid,
                     first_move=self._moves)
        chain.update(self._chain_extra())
        self._batch = ChainBatch([self.calc.model], [chain])
        self._batch_occ = occ.copy()
        self._batch_eci = eci
        return self._batch

    def _sync_back(self):
        """Write the chain's lattice back into the calculator and atoms."""
        occ = self._batch.occupation(0)
        self.calc.updater.set_occupation(occ)
        self.calc.cf = self.calc.updater.get_cf()
        self._batch_occ = occ.copy()
        self.current_energy = self.calc.get_energy()

    def _check_symbols(self):
        """Canonical swaps need two species (montecarlo.py:297-319)."""
        present = set(a.symbol for a in self.atoms)
        if len(present) < 2:
            raise ValueError("At least two different species are needed for swap moves")

    # ------------------------------------------------------------ interface
    def reset(self):
        """Zero the averages (montecarlo.py:349-364)."""
        self.current_step = 0
        self.num_accepted = 0
        if self._batch is not None:
            self._batch.reset_stats(0)

    def update_current_energy(self):
        self.current_energy = self.calc.get_energy()

    def set_symbols(self, symbs):
        """Replace every symbol (montecarlo.py:284-295)."""
        self.calc.set_symbols(symbs)
        self.update_current_energy()

    def _burn_in(self, equil_params):
        steps = int(equil_params.get("steps", 10 * len(self.atoms)))
        if steps > 0:
            batch = self._ensure_batch()
            batch.run(steps)
            self._moves += steps

    def runMC(self, mode="fixed", steps=10, verbose=False, equil=True, equil_params=None, prec=0.01,
              prec_confidence=0.05):
        """Run ``steps`` trial moves on the GPU (montecarlo.py:732-844).

        ``equil=True`` first runs a burn-in of ``equil_params["steps"]``
        moves (default 10 sweeps) whose samples are discarded; the reference
        instead equilibrates adaptively.  Only ``mode="fixed"`` is available.
        """
        if mode not in ("fixed", "prec"):
            raise ValueError("Mode has to be one of ['fixed', 'prec']")
        if mode == "prec":
            raise NotImplementedError("precision mode needs the correlation-time estimate, not on the GPU path")
        self._check_symbols()
        self.update_current_energy()
        equil_params = dict(equil_params or {})
        if equil:
            self._burn_in(equil_params)
        batch = self._ensure_batch()
        batch.reset_stats(0)
        stats0 = batch.stats()
        steps = int(steps)
        batch.run(steps)
        self._moves += steps
        stats = batch.stats()
        self.current_step = steps
        self.num_accepted = int(stats["accepted"][0] - stats0["accepted"][0])
        self._sync_back()
        return [self.current_energy]

    def count_atoms(self):
        counts = {s: 0 for s in self.symbols}
        for a in self.atoms:
            counts[a.symbol] = counts.get(a.symbol, 0) + 1
        return counts

    def get_thermodynamic(self):
        """Averages since the last reset (montecarlo.py:861-888)."""
        th = self._batch.thermo(0) if self._batch is not None else None
        q = {}
        if th is not None and th["n_samples"] > 0:
            q["energy"] = th["mean_energy"]
            q["heat_capacity"] = th["var_energy"] / (KB_EV * self._T ** 2)
            q["energy_var"] = th["var_energy"]
        else:
            q["energy"] = self.current_energy
            q["heat_capacity"] = 0.0
            q["energy_var"] = 0.0
        q["temperature"] = self._T
        q["accept_rate"] = self.num_accepted / self.current_step if self.current_step else 0.0
        n = len(self.atoms)
        for key, value in self.count_atoms().items():
            q["{}_conc".format(key)] = float(value) / n
        return q


class SGCMonteCarlo(Montecarlo):
    """Semi-grand-canonical sampling: single-site flips at fixed chemical
    potentials (sgc_montecarlo.py:12-76).  ``chem_potential`` maps singlet
    names to values in eV, e.g. {"c1_0": -0.1}; they are subtracted from the
    singlet ECIs of the chain (sgc_montecarlo.py:239-261)."""

    kind = SGC

    def __init__(self, atoms, temp, indeces=None, symbols=None, logfile="", plot_debug=False, min_acc_rate=0.0,
                 recycle_waste=False, seed=None, uid=0):
        Montecarlo.__init__(self, atoms, temp, indeces=indeces, logfile=logfile, plot_debug=plot_debug,
                            min_acc_rate=min_acc_rate, recycle_waste=recycle_waste, seed=seed, uid=uid)
        if symbols is not None:
            self.symbols = list(symbols)
        if len(self.symbols) <= 1:
            raise ValueError("At least 2 symbols have to be specified")
        if sorted(self.symbols) != sorted(self.calc.model.species):
            raise ValueError("SGC flips run over every species of the model: %s" % self.calc.model.species)
        self.name = "SGCMonteCarlo"
        self._chemical_potential = None

    @property
    def chemical_potential(self):
        return self._chemical_potential

    @chemical_potential.setter
    def chemical_potential(self, chem_pot):
        names = self.calc.model.names
        for key in chem_pot:
            if key not in names or not key.startswith("c1_"):
                raise ValueError("A chemical potential that is currently not tracked is added: %s" % key)
        self._chemical_potential = dict(chem_pot)
        if self._batch is not None:
            self._batch.close()
            self._batch = None

    def _chain_extra(self):
        return {"chem_pot": self._chemical_potential}

    def _check_symbols(self):
        pass

    def runMC(self, mode="fixed", steps=10, verbose=False, chem_potential=None, equil=True, equil_params=None,
              prec_confidence=0.05, prec=0.01):
        """sgc_montecarlo.py:336-374."""
        if chem_potential is None and self._chemical_potential is None:
            raise ValueError("No chemicalpotentials given. Has to be dictionary of the form "
                             "{'c1_1': -0.1, 'c1_2': 0.05}")
        if chem_potential is not None:
            self.chemical_potential = chem_potential
        return Montecarlo.runMC(self, mode=mode, steps=steps, verbose=verbose, equil=equil,
                                equil_params=equil_params, prec=prec, prec_confidence=prec_confidence)

    def _sync_back(self):
        # the calculator holds the plain ECIs; only the chain has mu folded in
        occ = self._batch.occupation(0)
        self.calc.updater.set_occupation(occ)
        self.calc.cf = self.calc.updater.get_cf()
        self._batch_occ = occ.copy()
        self.current_energy = self.calc.get_energy()

    def singlet2composition(self, avg_singlets):
        """sgc_montecarlo.py:376-396."""
        bf = self.calc.BC.basis_functions
        n = len(self.symbols)
        matrix = np.zeros((n, n))
        index = {s: i for i, s in enumerate(self.symbols)}
        for i, b in enumerate(bf):
            for s, col in index.items():
                matrix[i, col] = b[s]
        matrix[-1, :] = 1.0
        rhs = np.zeros(n)
        rhs[:-1] = avg_singlets
        rhs[-1] = 1.0
        x = np.linalg.solve(matrix, rhs)
        return {s + "_conc": float(x[i]) for s, i in index.items()}

    def get_thermodynamic(self, reset_ecis=True):
        """sgc_montecarlo.py:398-448."""
        th = self._batch.thermo(0)
        N = len(self.atoms)
        names = [n for n in self.calc.model.names if n.startswith("c1_")]
        mu = np.zeros(len(names))
        for k, name in enumerate(names):
            mu[k] = (self._chemical_potential or {}).get(name, 0.0)
        q = {}
        q["sgc_energy"] = th["mean_energy"]
        q["sgc_heat_capacity"] = th["var_energy"] / (KB_EV * self._T ** 2)
        q["energy"] = th["mean_energy"] + float(np.dot(mu, th["singlet"])) * N
        q["temperature"] = self._T
        q["n_mc_steps"] = th["n_samples"]
        for k, name in enumerate(names):
            q["singlet_{}".format(name)] = float(th["singlet"][k])
            q["var_singlet_{}".format(name)] = float(th["var_singlet"][k])
            q["mu_{}".format(name)] = float(mu[k])
        try:
            q.update(self.singlet2composition(th["singlet"]))
        except np.linalg.LinAlgError:
            pass
        return q
