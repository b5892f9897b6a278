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

``runMC`` follows the reference's protocol: with ``equil=True`` a
correlation-time estimate from a 1000-move energy trace
(montecarlo.py:461-526) and adaptive equilibration over windows of 10 N moves
compared by a z-test on their mean energies (and, in the SGC ensemble, on
their mean singlets; montecarlo.py:541-697, sgc_montecarlo.py:172-210);
``mode="prec"`` stops once the variance of the mean energy (SGC: of the mean
singlets) is below (prec / z_{1-confidence})^2, checked every ten
correlation times (montecarlo.py:699-716, 790-844).  Each window is one GPU
launch whose sums come from the device; the trace is read move by move.
The energy-bias probe of the reference (montecarlo.py:178-217) keeps its
squared sums well conditioned; the device sums here are taken relative to
the energy at the last reset for the same purpose.

Constraints and bias potentials (``add_constraint``, ``add_bias``; see
cemc_amd.constraints) and waste recycling (``recycle_waste=True``:
heat-bath acceptance, each move's sample weighting the current and the
proposed state, montecarlo.py:805-811, 937-948) run on the device, one wave
and one move at a time.  Observers (``attach``) are called between launches:
the run pauses at their intervals and each sees the sites changed since its
previous call.
"""
from __future__ import annotations

import logging

import numpy as np

from .batch import CANONICAL, SGC, ChainBatch
from .constraints import hooks_for
from .lattice import KB_EV


def _fresh_seed():
    return int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])


class CanNotFindLegalMoveError(Exception):
    """No proposal of a move passed the constraints (montecarlo.py:986-989)."""


class DidNotReachEquillibriumError(Exception):
    """montecarlo.py:27."""


def _norm_ppf(p):
    from scipy.stats import norm
    return float(norm.ppf(p))


def correlation_time(energies):
    """Correlation time of an energy series (montecarlo.py:461-526, after
    van de Walle and Asta 2002): the autocorrelation function normalised by
    N * var drops below 1/2 first at lag i; then rho = 2^(-1/i) and
    tau = -1 / ln(rho).  Returns the reference's info dictionary."""
    e = np.asarray(energies, dtype=np.float64)
    n = len(e)
    info = {"correlation_time_found": False, "correlation_time": 0.0, "msg": ""}
    dev = e - e.mean()
    var = np.var(dev)
    if var == 0.0:
        info["msg"] = "Zero variance leads to infinite correlation time"
        info["correlation_time_found"] = True
        info["correlation_time"] = n
        return info
    ac = np.correlate(dev, dev, mode="full")
    ac = ac[int(len(ac) / 2):] / (n * var)
    if np.min(ac) > 0.5:
        info["msg"] = "Window is too short. Add more samples"
        info["correlation_time"] = n
        return info
    indx = 0
    for i in range(len(ac)):
        if ac[i] < 0.5:
            indx = i
            break
    rho = 2.0 ** (-1.0 / indx)
    info["correlation_time"] = -1.0 / np.log(rho)
    info["correlation_time_found"] = True
    return info


def variance_of_mean(var, n, corr_info):
    """Variance of an average of n correlated samples with variance var
    (montecarlo.py:417-447): var / n without a correlation time, else
    2 var tau / n with tau at least 1."""
    var = np.abs(var)
    if corr_info is None or not corr_info["correlation_time_found"]:
        return var / n
    tau = max(corr_info["correlation_time"], 1.0)
    return 2.0 * var * tau / n


def prec_check_steps(check_every, total):
    """Step counts after which the reference's prec mode tests convergence
    (the first step past each multiple of check_every, montecarlo.py:790-838),
    up to total."""
    out = []
    nxt = check_every
    while True:
        s = int(np.floor(nxt)) + 1
        if s >= total:
            break
        out.append(s)
        nxt += check_every
    return out


# temperature at which every trial move is accepted (see Montecarlo._forced_first)
_ACCEPT_ALL_T = 1e300


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
        self.correlation_info = None
        self.converged = False
        self.recycle_waste = bool(recycle_waste)
        self.constraints = []
        self.bias_potentials = []
        self.observers = []
        self.logger = logging.getLogger("cemc_amd.mcmc.%x" % id(self))
        if logfile:
            handler = logging.FileHandler(logfile)
            handler.setFormatter(logging.Formatter("%(asctime)s %(message)s"))
            self.logger.addHandler(handler)
            self.logger.setLevel(logging.INFO)
        self.max_allowed_constraint_pass_attempts = int(max_constraint_attempts)
        self.accept_first_trial_move_after_reset = bool(accept_first_trial_move_after_reset)
        if self.accept_first_trial_move_after_reset and self.recycle_waste:
            raise NotImplementedError("accept_first_trial_move_after_reset with recycle_waste")
        self.is_first = self.accept_first_trial_move_after_reset
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

    @property
    def linear_vib_correction(self):
        return self.calc.linear_vib_correction

    @linear_vib_correction.setter
    def linear_vib_correction(self, linvib):
        """montecarlo.py:147-154.  The averages leave the vibrational energy
        out exactly when it depends on the chain's sums only: terms on the
        empty cluster and the singlets (in the semi-grand-canonical ensemble,
        on at most one singlet, whose variance and covariance with the energy
        the device tracks)."""
        names = self.calc.model.names
        singl = [k for k in linvib.eci_per_kbT if k.startswith("c1_")]
        for key in linvib.eci_per_kbT:
            if key not in names:
                raise KeyError("The cluster {} is not in the original ECIs!".format(key))
            if key != "c0" and not key.startswith("c1_"):
                raise NotImplementedError("vibrational terms on cluster %s: the samplers average the energy "
                                          "without vibrations for the empty cluster and the singlets only" % key)
        if self.kind == SGC and len(singl) > 1:
            raise NotImplementedError("vibrational terms on more than one singlet in the semi-grand-canonical "
                                      "ensemble")
        self.calc.linear_vib_correction = linvib

    def _include_vib(self):
        """montecarlo.py:366-370."""
        self.calc.include_linvib_in_ecis(self._T)

    def current_energy_without_vib(self):
        """montecarlo.py:452-459."""
        return self.current_energy - self.calc.vib_energy(self._T) * len(self.atoms)

    def _vib_shift(self, th):
        """Mean and variance corrections that take the vibrational energy
        N kB T (v_c0 + sum_d v_d c1_d) out of the chain's energy averages:
        (mean of that energy, variance change)."""
        lv = self.calc.linear_vib_correction
        if lv is None or not lv.eci_per_kbT:
            return 0.0, 0.0
        from .linear_vib import KB_VIB
        n = len(self.atoms)
        names = [k for k in self.calc.model.names if k.startswith("c1_")]
        scale = n * KB_VIB * self._T
        mean = scale * lv.eci_per_kbT.get("c0", 0.0)
        dvar = 0.0
        for d, name in enumerate(names):
            a = scale * lv.eci_per_kbT.get(name, 0.0)
            if a == 0.0:
                continue
            mean += a * th["singlet"][d]
            if self.kind == SGC:
                dvar += a * a * th["var_singlet"][d] - 2.0 * a * th["cov_singlet_energy"][d]
        return mean, dvar

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
        if self.constraints or self.bias_potentials or self.recycle_waste:
            self._batch.set_hooks(0, **hooks_for(self.calc.BC, self.calc.model.species, self.constraints,
                                                 self.bias_potentials, self.max_allowed_constraint_pass_attempts,
                                                 recycle_waste=self.recycle_waste))
        return self._batch

    def _sync_back(self):
        """Write the chain's lattice back into the calculator and atoms."""
        occ = self._batch.occupation(0)
        self.calc.updater.set_occupation(occ)
        self.calc.cf = self.calc.updater.get_cf()
        self._batch_occ = occ.copy()
        self.current_energy = self.calc.get_energy() + self._bias_energy()

    def _bias_energy(self):
        """Bias of the current state; part of current_energy as in the
        reference (montecarlo.py:279-282)."""
        if not self.bias_potentials or self._batch is None:
            return 0.0
        return self._batch.hooks_state(0)[2]

    def add_constraint(self, constraint):
        """Add a constraint checked per trial move on the device
        (montecarlo.py:386-392); see cemc_amd.constraints."""
        if not hasattr(constraint, "device_hooks"):
            raise TypeError("constraints must be cemc_amd.constraints classes (checked on the device)")
        self.constraints.append(constraint)
        self._drop_batch()

    def log(self, msg, mode="info"):
        """montecarlo.py:321-332: to the sampler's logger (a file when
        ``logfile`` was given, else the process's logging setup)."""
        if mode not in ("info", "warning"):
            raise ValueError("Mode has to be one of ['info', 'warning']")
        getattr(self.logger, mode)(msg)

    def _no_constraint_violations(self, system_changes):
        """Host check of changes against every constraint
        (montecarlo.py:334-347); trial moves are checked on the device."""
        return all(c(system_changes) for c in self.constraints)

    def insert_symbol(self, symb, indices):
        """Put ``symb`` on the given sites (montecarlo.py:218-233)."""
        for indx in indices:
            change = (int(indx), self.atoms[int(indx)].symbol, symb)
            if not self._no_constraint_violations([change]):
                raise ValueError("The indices given results in an update that violate one or more constraints!")
            self.calc.update_cf(change)
        self.calc.clear_history()
        self.calc.cf = self.calc.updater.get_cf()
        self.update_current_energy()

    def insert_symbol_random_places(self, symbol, num=1, swap_symbs=()):
        """Put ``symbol`` on ``num`` random sites holding one of
        ``swap_symbs`` (all symbols present when empty), skipping changes a
        constraint forbids, with at most 10 N attempts
        (montecarlo.py:235-270).  Sites are drawn from the sampler's seed."""
        rng = np.random.default_rng(None if self.seed is None else [int(self.seed), int(self.uid), 0x1a5e])
        symbols = [a.symbol for a in self.atoms]
        swap_symbs = list(swap_symbs) or sorted(set(symbols))
        inserted = attempts = 0
        max_attempts = 10 * len(self.atoms)
        while inserted < num and attempts < max_attempts:
            attempts += 1
            old = swap_symbs[int(rng.integers(len(swap_symbs)))]
            if old == symbol:
                continue
            sites = [i for i, s in enumerate(symbols) if s == old]
            if not sites:
                continue
            indx = sites[int(rng.integers(len(sites)))]
            change = (indx, old, symbol)
            if not self._no_constraint_violations([change]):
                continue
            self.calc.update_cf(change)
            symbols[indx] = symbol
            inserted += 1
        self.calc.clear_history()
        self.calc.cf = self.calc.updater.get_cf()
        self.update_current_energy()
        if attempts == max_attempts:
            raise RuntimeError("Could insert {} {} atoms!".format(num, symbol))

    def attach(self, obs, interval=1):
        """Call ``obs(system_changes)`` every ``interval`` moves of a
        production run (montecarlo.py:403-415, 1032-1037).  The run then
        pauses at those moves: the lattice is written back to the atoms and
        the calculator first, and ``system_changes`` lists the sites changed
        since this observer's previous call as (index, old symbol, new
        symbol); for interval 1 that is the move's changes when it was
        accepted and an empty list when it was rejected."""
        if not callable(obs):
            raise ValueError("The observer has to be a callable class!")
        self.observers.append((int(interval), obs))

    def _advance_observed(self, begin, end):
        """Moves begin..end of the production run, pausing for observers."""
        if not self.observers:
            self._advance(end - begin)
            return
        species = self.calc.model.species
        self._ensure_batch()
        last = [self._batch_occ.copy() for _ in self.observers]
        cur = begin
        while cur < end:
            nxt = min(min((cur // i + 1) * i for i, _ in self.observers), end)
            self._advance(nxt - cur)
            cur = nxt
            due = [k for k, (i, _) in enumerate(self.observers) if cur % i == 0]
            if not due:
                continue
            self._sync_back()
            occ = self._batch_occ
            self.current_step = cur
            for k in due:
                prev = last[k]
                changes = [(int(j), species[prev[j]], species[occ[j]]) for j in np.nonzero(prev != occ)[0]]
                last[k] = occ.copy()
                self.observers[k][1](changes)

    def add_bias(self, potential):
        """Add a bias potential (montecarlo.py:394-401); see
        cemc_amd.constraints.ConcentrationBias."""
        if not hasattr(potential, "device_table"):
            raise TypeError("potential has to be a cemc_amd.constraints bias potential")
        self.bias_potentials.append(potential)
        self._drop_batch()

    def _drop_batch(self):
        if self._batch is not None:
            self._sync_back()
            self._batch.close()
            self._batch = None

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
        if self.accept_first_trial_move_after_reset:
            self.is_first = True

    def update_current_energy(self):
        self.current_energy = self.calc.get_energy()

    def set_symbols(self, symbs):
        """Replace every symbol (montecarlo.py:284-295)."""
        self.calc.set_symbols(symbs)
        self.update_current_energy()
        if self.accept_first_trial_move_after_reset:
            self.is_first = True

    def _forced_first(self, batch, n):
        """With accept_first_trial_move_after_reset, the first move after a
        reset is accepted whatever its energy (montecarlo.py:932-935): it
        runs at 1e300 K, where |dE / kT| < 2^-53 for any energy change, so
        exp(-dE/kT) rounds to 1 and every uniform u < 1 accepts.  Returns
        the moves left."""
        if n <= 0 or not self.is_first:
            return n
        self.is_first = False
        batch.set_temperature(0, _ACCEPT_ALL_T)
        try:
            batch.run(1)
        finally:
            batch.set_temperature(0, self._T)
        self._moves += 1
        return n - 1

    def _advance(self, n):
        batch = self._ensure_batch()
        n = self._forced_first(batch, int(n))
        try:
            batch.run(int(n))
        except RuntimeError as exc:
            if "Did not manage to produce a trial move" not in str(exc):
                raise
            self._moves = int(batch.stats()["moves"][0])
            self._sync_back()
            raise CanNotFindLegalMoveError(str(exc)) from None
        self._moves += int(n)
        return batch

    def _energy_trace(self, n):
        """Energy after each of the next n moves (one launch per move)."""
        batch = self._ensure_batch()
        out = np.empty(int(n))
        for i in range(int(n)):
            if self.is_first:
                self._forced_first(batch, 1)
            else:
                batch.run(1)
                self._moves += 1
            out[i] = batch.energies()[0]
        return out

    def _estimate_correlation_time(self, window_length=1000, restart=False):
        """montecarlo.py:461-526."""
        if restart or not hasattr(self, "corrtime_energies"):
            self.corrtime_energies = []
        self.corrtime_energies.extend(self._energy_trace(window_length).tolist())
        self.correlation_info = correlation_time(self.corrtime_energies)
        return self.correlation_info

    def _window(self, n):
        """Reset the sums, run n moves, return the device averages."""
        batch = self._ensure_batch()
        batch.reset_stats(0)
        self._advance(n)
        return batch.thermo(0)

    def _get_var_average_energy(self, th=None):
        """montecarlo.py:417-447."""
        th = th if th is not None else self._batch.thermo(0)
        return variance_of_mean(th["var_energy"], max(th["n_samples"], 1), self.correlation_info)

    def _composition_reached_equillibrium(self, prev_composition, var_prev, th, confidence_level=0.05):
        """Fixed composition: always (montecarlo.py:530-539)."""
        return True, prev_composition, var_prev, 0.0

    def _equillibriate(self, window_length="auto", confidence_level=0.05, maxiter=1000, mode="stat_equiv"):
        """montecarlo.py:541-697: windows of ``window_length`` moves (10 N
        for "auto") until the mean energies of two consecutive windows differ
        by less than the confidence band of a z-test (and, in the SGC
        ensemble, the mean singlets as well); mode "fixed" runs one window."""
        if mode not in ("stat_equiv", "fixed"):
            raise ValueError("Equilibration mode has to be one of ['stat_equiv', 'fixed']")
        if window_length == "auto":
            window_length = 10 * len(self.atoms)
        window_length = int(window_length)
        self.reset()
        if mode == "fixed":
            self._advance(window_length)
            return 0
        lo, hi = _norm_ppf(confidence_level), _norm_ppf(1.0 - confidence_level)
        e_prev = var_prev = None
        composition, var_comp = [], []
        for it in range(maxiter):
            th = self._window(window_length)
            e_new = th["mean_energy"]
            var_new = self._get_var_average_energy(th)
            comp_conv, composition, var_comp, _ = self._composition_reached_equillibrium(
                composition, var_comp, th, confidence_level=confidence_level)
            if e_prev is None:
                e_prev, var_prev = e_new, var_new
                continue
            var_diff = var_new + var_prev
            if var_diff < 1e-6:
                z = 0.0
                comp_conv = True
            else:
                z = (e_new - e_prev) / np.sqrt(var_diff)
            if lo < z < hi and comp_conv:
                self.reset()
                return (it + 2) * window_length
            e_prev, var_prev = e_new, var_new
        raise DidNotReachEquillibriumError("Did not manage to reach equillibrium!")

    def _has_converged_prec_mode(self, prec=0.01, confidence_level=0.05, th=None):
        """montecarlo.py:699-716."""
        percentile = _norm_ppf(1.0 - confidence_level)
        return self._get_var_average_energy(th) < (prec / percentile) ** 2

    def runMC(self, mode="fixed", steps=10, verbose=False, equil=True, equil_params=None, prec=0.01,
              prec_confidence=0.05):
        """Run up to ``steps`` trial moves on the GPU (montecarlo.py:732-844).

        ``equil=True``: correlation time from a 1000-move trace, then
        adaptive equilibration (``equil_params`` as ``_equillibriate``).
        ``mode="prec"``: the correlation time is estimated (longer traces
        until found) and the run stops at the first convergence check, every
        ten correlation times, that meets ``prec``.  The averages of
        ``get_thermodynamic`` cover the production run only.
        """
        if mode not in ("fixed", "prec"):
            raise ValueError("Mode has to be one of ['fixed', 'prec']")
        self._check_symbols()
        self._include_vib()
        self.update_current_energy()
        self.correlation_info = None
        if equil:
            res = self._estimate_correlation_time(restart=True)
            if not res["correlation_time_found"]:
                res["correlation_time"] = 1000
                res["correlation_time_found"] = True
            self._equillibriate(**dict(equil_params or {}))
        checks = []
        if mode == "prec":
            res = self._estimate_correlation_time(restart=True)
            while not res["correlation_time_found"]:
                res = self._estimate_correlation_time()
            checks = prec_check_steps(10 * self.correlation_info["correlation_time"], int(steps))
        batch = self._ensure_batch()
        # the reference calls reset() right before the production moves
        # (montecarlo.py:796-800), which also re-arms the forced first move
        batch.reset_stats(0)
        if self.accept_first_trial_move_after_reset:
            self.is_first = True
        stats0 = batch.stats()
        steps = int(steps)
        done = 0
        self.converged = False
        for s in checks + [steps]:
            self._advance_observed(done, s)
            done = s
            if s < steps and self._has_converged_prec_mode(prec, prec_confidence, batch.thermo(0)):
                self.converged = True
                break
        stats = batch.stats()
        self.current_step = done
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
            vib_mean, vib_dvar = self._vib_shift(th)
            th = dict(th, mean_energy=th["mean_energy"] - vib_mean, var_energy=th["var_energy"] + vib_dvar)
            q["energy"] = th["mean_energy"]
            q["heat_capacity"] = th["var_energy"] / (KB_EV * self._T ** 2)
            q["energy_var"] = th["var_energy"]
            q["energy_std"] = float(np.sqrt(max(self._get_var_average_energy(th), 0.0)))
        else:
            q["energy"] = self.current_energy
            q["heat_capacity"] = 0.0
            q["energy_var"] = 0.0
            q["energy_std"] = 0.0
        q["temperature"] = self._T
        q["accept_rate"] = self.num_accepted / self.current_step if self.current_step else 0.0
        n = len(self.atoms)
        for key, value in self.count_atoms().items():
            q["{}_conc".format(key)] = float(value) / n
        q.update(self.meta_info)
        q.update(self._observer_averages())
        return q

    def save(self, fname):
        """Write the sampler's state as JSON (montecarlo.py:1040-1057 pickles
        it): the calculator's backup, temperature, seed, uid, the number of
        moves made (a loaded sampler continues the same move stream), the
        chemical potentials and the vibrational correction.  Constraints,
        bias potentials and observers are not stored and are added again
        after ``load``, as the reference says of those it cannot pickle."""
        import json
        lv = self.calc.linear_vib_correction
        state = {"class": type(self).__name__, "calc": self.calc.backup_dict(), "temperature": self._T,
                 "seed": self.seed, "uid": self.uid, "moves": int(self._moves),
                 "recycle_waste": self.recycle_waste,
                 "accept_first_trial_move_after_reset": self.accept_first_trial_move_after_reset,
                 "is_first": bool(self.is_first),
                 "max_constraint_attempts": self.max_allowed_constraint_pass_attempts,
                 "chem_pot": getattr(self, "_chemical_potential", None),
                 "linear_vib": None if lv is None else {"eci_per_kbT": lv.eci_per_kbT, "included": lv.vibs_included,
                                                        "temperature": lv.temperature}}
        with open(fname, "w") as f:
            json.dump(state, f, indent=2)

    @staticmethod
    def load(fname):
        """Rebuild a sampler written by ``save`` (montecarlo.py:1059-1075)."""
        import json
        from .calculator import CE
        from .linear_vib import LinearVibCorrection
        with open(fname) as f:
            st = json.load(f)
        calc = CE.load_from_dict(st["calc"])
        kw = dict(seed=st["seed"], uid=st["uid"], recycle_waste=st["recycle_waste"],
                  max_constraint_attempts=st["max_constraint_attempts"])
        if st["class"] == "SGCMonteCarlo":
            # the SGC sampler has no forced first move (sgc_montecarlo.py:31-36)
            mc = SGCMonteCarlo(calc.atoms, st["temperature"], symbols=list(calc.model.species), **kw)
            if st["chem_pot"] is not None:
                mc.chemical_potential = st["chem_pot"]
        else:
            mc = Montecarlo(calc.atoms, st["temperature"],
                            accept_first_trial_move_after_reset=st.get("accept_first_trial_move_after_reset", False),
                            **kw)
        if st["linear_vib"] is not None:
            lv = LinearVibCorrection(st["linear_vib"]["eci_per_kbT"])
            lv.vibs_included = bool(st["linear_vib"]["included"])
            lv.temperature = float(st["linear_vib"]["temperature"])
            calc._linear_vib_correction = lv
            calc.updater.add_linear_vib_correction(lv.eci_per_kbT)
        mc._moves = int(st["moves"])
        mc.is_first = bool(st.get("is_first", mc.is_first))
        return mc

    def _observer_averages(self):
        """Averages of the attached observers that report them
        (montecarlo.py:885-887)."""
        out = {}
        for _, obs in self.observers:
            if hasattr(obs, "get_averages"):
                out.update(obs.get_averages())
        return out

    @property
    def meta_info(self):
        """Timestamp and Python version (montecarlo.py:846-859)."""
        import datetime
        import sys
        v = sys.version_info
        return {"timestamp": datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S"),
                "python_version": "{}.{}.{}".format(v.major, v.minor, v.micro)}


class SGCMonteCarlo(Montecarlo):
    """Semi-grand-canonical sampling: single-site flips at fixed chemical
    potentials (sgc_montecarlo.py:12-76).  ``chem_potential`` maps singlet
    names to values in eV, e.g. {"c1_0": -0.1}; they are subtracted from the
    singlet ECIs of the chain (sgc_montecarlo.py:239-261)."""

    kind = SGC

    def __init__(self, atoms, temp, indeces=None, symbols=None, logfile="", plot_debug=False, min_acc_rate=0.0,
                 recycle_waste=False, seed=None, uid=0, max_constraint_attempts=10000):
        Montecarlo.__init__(self, atoms, temp, indeces=indeces, logfile=logfile, plot_debug=plot_debug,
                            min_acc_rate=min_acc_rate, recycle_waste=recycle_waste,
                            max_constraint_attempts=max_constraint_attempts, seed=seed, uid=uid)
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

    def _get_var_average_singlets(self, th):
        """sgc_montecarlo.py:94-130."""
        return variance_of_mean(np.asarray(th["var_singlet"]), max(th["n_samples"], 1), self.correlation_info)

    def _composition_reached_equillibrium(self, prev_composition, var_prev, th, confidence_level=0.05):
        """sgc_montecarlo.py:172-210: z-test on the largest change of the
        mean singlets between windows."""
        lo, hi = _norm_ppf(confidence_level), _norm_ppf(1.0 - confidence_level)
        singlets = np.asarray(th["singlet"], dtype=np.float64)
        var_n = np.maximum(self._get_var_average_singlets(th), 0.0)
        if len(prev_composition) != len(singlets):
            return False, singlets, var_n, 0.0
        diff = singlets - np.asarray(prev_composition)
        var_diff = var_n + np.asarray(var_prev)
        pos = var_diff > 0.0
        if not np.any(pos):
            return True, singlets, var_n, 0.0
        z = float(np.max(np.abs(diff[pos]) / np.sqrt(var_diff[pos])))
        return lo < z < hi, singlets, var_n, z

    def _has_converged_prec_mode(self, prec=0.01, confidence_level=0.05, th=None):
        """sgc_montecarlo.py:132-155: the reference decides on the singlets
        alone (its energy test is computed and not used)."""
        th = th if th is not None else self._batch.thermo(0)
        percentile = _norm_ppf(1.0 - confidence_level)
        return float(np.max(self._get_var_average_singlets(th))) < (prec / percentile) ** 2

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
        self.current_energy = self.calc.get_energy() + self._bias_energy()

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
        vib_mean, vib_dvar = self._vib_shift(th)
        th = dict(th, mean_energy=th["mean_energy"] - vib_mean, var_energy=th["var_energy"] + vib_dvar)
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
        q.update(self.meta_info)
        try:
            q.update(self.singlet2composition(th["singlet"]))
        except np.linalg.LinAlgError:
            pass
        return q

    def reset_ecis(self):
        """sgc_montecarlo.py:276-281.  The calculator here always holds the
        plain ECIs (the chemical potentials live in the chain's parameters
        only), so there is nothing to undo; ``get_thermodynamic``'s
        ``reset_ecis`` argument is accepted for the same reason."""
        return None
