"""Cluster-expansion calculator on the GPU: the reference's ``CE`` class
(cemc/ce_calculator.py:136-470) over ``cemc_amd.updater.PyCEUpdater``.

Same constructor and methods; the correlation functions, the history and the
occupation live on the GPU.  The reference derives from ASE's ``Calculator``;
this class keeps the parts of that contract its callers use (``results``,
``get_potential_energy``, ``atoms.set_calculator``) without importing ASE.
"""
from __future__ import annotations

import numpy as np

from .linear_vib import LinearVibCorrection
from .updater import PyCEUpdater


class SelfInteractionError(ValueError):
    """A site appears twice in one cluster (cemc/ce_calculator.py:20, 152-157)."""


def _has_self_interaction(cluster_info):
    """True when a cluster lists its own reference site or one site twice
    (the condition CE._has_self_interaction tests, ce_calculator.py:596-608)."""
    for info in cluster_info:
        for cl in info.values():
            if int(cl.get("size", 0)) < 2:
                continue
            ref = int(cl["ref_indx"])
            for sub in cl["indices"]:
                sub = [int(v) for v in sub]
                if ref in sub or len(set(sub)) != len(sub):
                    return True
    return False


class CE:
    """Energy of an atoms object from a cluster expansion.

    :param atoms: object with ``len``, ``atoms[i].symbol`` (read and write)
        and ``set_calculator``
    :param BC: settings (the attribute contract of ``CEUpdater::init``,
        cpp/src/ce_updater.cpp:32-234)
    :param dict eci: effective cluster interactions by CF name
    :param initial_cf: correlation functions by name; None computes them on
        the GPU from the occupation (the reference uses CLEASE's
        ``CorrFunction`` here, ce_calculator.py:164-172)
    """

    implemented_properties = ["energy"]

    def __init__(self, atoms, BC, eci=None, initial_cf=None):
        self.BC = BC
        if _has_self_interaction(BC.cluster_info):
            raise SelfInteractionError(
                "The simulation cell is so small that the same site is present multiple times within one "
                "cluster. Increase the size of the simulation cell.")
        eci = dict(eci) if eci is not None else {}
        # an ECI for the empty cluster is always present (ce_calculator.py:152-158)
        if "c0" not in eci:
            eci["c0"] = 0.0
            if initial_cf is not None:
                initial_cf = dict(initial_cf)
                initial_cf["c0"] = 1.0
        self.eci = eci
        self.atoms = atoms
        self._check_trans_mat_dimensions()
        self.results = {}
        self.updater = PyCEUpdater(atoms, BC, initial_cf, eci)
        self.cf = self.updater.get_cf()
        self.clear_history = self.updater.clear_history
        self.undo_changes = self.updater.undo_changes
        self.update_cf = self.updater.update_cf
        if hasattr(atoms, "set_calculator"):
            atoms.set_calculator(self)
        self._linear_vib_correction = None

    @property
    def model(self):
        """The device model (lattice, clusters, ECIs) behind this calculator."""
        return self.updater.model

    def _check_trans_mat_dimensions(self):
        tm = self.BC.trans_matrix
        n_sites = tm.shape[0] if isinstance(tm, np.ndarray) else len(tm)
        if len(self.atoms) != n_sites:
            raise ValueError("The number of atoms and the dimension of the translation matrix is inconsistent\n"
                             "Num atoms: {}. Num row trans mat: {}".format(len(self.atoms), n_sites))

    def backup_dict(self):
        """Everything needed to rebuild the calculator in its current state
        (ce_calculator.py:532-546): CFs, symbols, settings arguments with
        their class, ECIs."""
        kwargs = dict(getattr(self.BC, "kwargs", {}))
        kwargs["classtype"] = type(self.BC).__name__
        return {"cf": dict(self.get_cf()), "symbols": [a.symbol for a in self.atoms],
                "setting_kwargs": kwargs, "eci": dict(self.eci)}

    def save(self, fname):
        """JSON file with backup_dict() (ce_calculator.py:548-557)."""
        import json
        with open(fname, "w") as outfile:
            json.dump(self.backup_dict(), outfile, indent=2, separators=(",", ": "))

    @staticmethod
    def load(fname):
        """Rebuild a calculator saved by save() (ce_calculator.py:562-571)."""
        import json
        with open(fname, "r") as infile:
            return CE.load_from_dict(json.load(infile))

    @staticmethod
    def load_from_dict(backup_data):
        """ce_calculator.py:573-592: settings from their arguments (see
        cemc_amd.lattice.settings_from_kwargs), atoms with the saved symbols
        site by site, the saved ECIs and CFs.  ECI names the settings do not
        define raise ValueError."""
        from .lattice import SimpleAtoms, settings_from_kwargs
        bc = settings_from_kwargs(backup_data["setting_kwargs"])
        names = set(bc.cf_names()) | {"c0"}
        unknown = sorted(set(backup_data["eci"]) - names)
        if unknown:
            raise ValueError("ECIs not defined by the settings: {}".format(unknown))
        symbols = list(backup_data["symbols"])
        if len(symbols) != bc.n_sites:
            raise ValueError("{} symbols for {} sites".format(len(symbols), bc.n_sites))
        atoms = SimpleAtoms(symbols)
        cf = backup_data.get("cf")
        cf = {k: v for k, v in cf.items() if k in names} if cf else None
        return CE(atoms, bc, eci=backup_data["eci"], initial_cf=cf)

    def __reduce__(self):
        return (CE.load_from_dict, (self.backup_dict(),))

    def copy(self):
        """A new calculator on a copy of the atoms with the current CFs."""
        atoms = self.atoms.copy()
        return CE(atoms, self.BC, eci=dict(self.eci), initial_cf=self.get_cf())

    def get_energy(self):
        return self.updater.get_energy()

    def get_potential_energy(self, atoms=None):
        return self.get_energy()

    def calculate(self, atoms, properties, system_changes):
        """Energy after applying ``system_changes`` = [(index, old, new), ...]
        (ce_calculator.py:345-364)."""
        energy = self.updater.calculate(system_changes)
        self.cf = self.updater.get_cf()
        self.results["energy"] = energy
        return energy

    def get_cf(self):
        return self.updater.get_cf()

    def update_ecis(self, new_ecis):
        """ECI update by name (ce_calculator.py:378-386); the device model's
        energy tables follow, so samplers built on it see the new values."""
        self.eci.update({k: float(v) for k, v in new_ecis.items()})
        self.updater.set_ecis(self.eci)
        self.model.set_ecis(self.eci)

    def get_singlets(self):
        return self.updater.get_singlets()

    def set_composition(self, comp):
        """Set concentrations, e.g. {"Mg": 0.2, "Al": 0.8}: the element with
        the largest fraction everywhere, then the others on consecutive sites
        (ce_calculator.py:387-423)."""
        tot = 0.0
        max_elem, max_conc = None, 0.0
        for key, conc in comp.items():
            tot += conc
            if conc > max_conc:
                max_elem, max_conc = key, conc
        if abs(tot - 1.0) > 1e-6:
            raise ValueError("The specified concentration does not sum to 1!")
        symbols = [max_elem] * len(self.atoms)
        start = 0
        for elem, conc in comp.items():
            if elem == max_elem:
                continue
            n_at = int(round(conc * len(self.atoms)))
            for i in range(start, start + n_at):
                symbols[i] = elem
            start += n_at
        self.set_symbols(symbols)

    def set_symbols(self, symbs):
        """Change every symbol at once (ce_calculator.py:436-448); CFs are
        recomputed on the GPU and the history cleared."""
        if len(symbs) != len(self.atoms):
            raise ValueError("Length of the symbols array has to match the length of the atoms object.!")
        self.updater.set_symbols(list(symbs))
        self.cf = self.updater.get_cf()

    def singlet2comp(self, singlets):
        """Concentrations from singlet values (ce_calculator.py:450-470)."""
        bfs = self.BC.basis_functions
        if len(singlets) != len(bfs):
            raise ValueError("The number singlet terms specified is different from the number of basis "
                             "functions\nGiven singlet terms: {}\nBasis functions: {}\n".format(singlets, bfs))
        species = sorted(self.BC.unique_elements)
        n = len(species)
        mat = np.zeros((n, n))
        rhs = np.zeros(n)
        names = sorted(singlets)
        for row, name in enumerate(names):
            d = int(name.rsplit("_", 1)[1])
            for col, s in enumerate(species):
                mat[row, col] = bfs[d][s]
            rhs[row] = singlets[name]
        mat[-1, :] = 1.0
        rhs[-1] = 1.0
        x = np.linalg.solve(mat, rhs)
        # round-off below zero is clipped as the reference does (ce_calculator.py:486-494)
        x[(x < 0.0) & (x > -1e-6)] = 0.0
        if np.any(x < 0.0) or np.any(x > 1.0):
            raise ValueError("Something strange happened when converting singlets to composition\n"
                             "Concentrations: {}".format(x))
        return {s: float(x[i]) for i, s in enumerate(species)}

    @property
    def linear_vib_correction(self):
        return self._linear_vib_correction

    @linear_vib_correction.setter
    def linear_vib_correction(self, linvib):
        """ce_calculator.py:269-287: a previous correction's terms leave the
        ECIs first; the updater gets the new coefficients (the ECIs themselves
        change only in ``include_linvib_in_ecis``)."""
        if not isinstance(linvib, LinearVibCorrection):
            raise TypeError("Linear vib correction has to be of type LinearVibCorrection!")
        if self._linear_vib_correction is not None:
            orig = self._linear_vib_correction.reset(self.eci)
            if orig is not None:
                self.eci = orig
            self.update_ecis(self.eci)
        self._linear_vib_correction = linvib
        self.updater.add_linear_vib_correction(linvib.eci_per_kbT)

    def include_linvib_in_ecis(self, T):
        """Add v_k kB T to the ECIs at temperature T (ce_calculator.py:289-301)."""
        if self._linear_vib_correction is None:
            return
        self._linear_vib_correction.include(self.eci, T)
        self.update_ecis(self.eci)

    def vib_energy(self, T):
        """Vibrational energy per site at the current CFs (ce_calculator.py:303-317)."""
        return self.updater.vib_energy(T)

    def get_full_cluster_names(self, cnames):
        """Names with their decoration suffix for cluster prefixes; one-body
        names are kept and listed first (ce_calculator.py:248-267)."""
        full_names = list(self.cf.keys())
        only_prefix = [name.rpartition("_")[0] for name in full_names]
        out = [name for name in cnames if name.startswith("c1")]
        for name in cnames:
            if not name.startswith("c1"):
                out.append(full_names[only_prefix.index(name)])
        return out

    def set_singlets(self, singlets):
        """Composition that gives these singlets, as close as the cell
        allows (ce_calculator.py:521-532)."""
        self.set_composition(self.singlet2comp(singlets))

    def set_num_threads(self, num_threads):
        """ce_calculator.py:610-617; the updater's work runs on GPU lanes."""
        self.updater.set_num_threads(num_threads)


def get_max_size_eci(eci):
    """Largest cluster size named in the ECIs: the digit after "c" in each
    name (ce_calculator.py:122-133)."""
    return max((int(key[1]) for key in eci), default=0)


def _max_tracked_size(bc_kwargs):
    """Largest cluster size the settings arguments track, or None when they
    name none (the reference reads ``max_cluster_size``; this package's
    settings key their cutoffs by cluster size)."""
    if "max_cluster_size" in bc_kwargs:
        return int(bc_kwargs["max_cluster_size"])
    for key in ("max_shell_by_size", "max_dia_by_size"):
        if bc_kwargs.get(key):
            return max(int(k) for k in bc_kwargs[key])
    return None


def template_symbols(bc):
    """Symbols of the settings' template cell: every site holds the first
    element of its sublattice (background sites their only element)."""
    sub = getattr(bc, "sublattice_of_site", None)
    if sub is None:
        return [bc.basis_elements[0][0]] * bc.n_sites
    return [bc.basis_elements[int(s)][0] for s in sub]


def get_atoms_with_ce_calc(small_bc, bc_kwargs, eci=None, size=(1, 1, 1), db_name="temp_db.db"):
    """Atoms of a large cell with a calculator whose initial correlation
    functions come from a small cell (ce_calculator.py:48-102).

    The template cell of ``small_bc`` (every site the first element of its
    sublattice) is evaluated on the GPU; the settings are rebuilt from
    ``bc_kwargs`` with ``size`` as the supercell, and the large template cell
    gets a calculator started from the small cell's correlation functions,
    which equal the large cell's because both are the same periodic crystal.
    ``bc_kwargs`` is not modified and ``db_name`` is accepted for the
    reference's signature (no database is written).

    :raises ValueError: the ECIs name a cluster size the settings do not
        track
    """
    from .lattice import SimpleAtoms, settings_from_kwargs
    del db_name
    eci = dict(eci) if eci is not None else {}
    tracked = _max_tracked_size(bc_kwargs)
    if tracked is not None and get_max_size_eci(eci) > tracked:
        raise ValueError("ECI specifies a cluster size larger than ClusterExpansionSetting tracks!")
    small_atoms = SimpleAtoms(template_symbols(small_bc))
    small_calc = CE(small_atoms, small_bc, eci)
    init_cf = small_calc.get_cf()
    small_calc.updater.close()
    kwargs = dict(bc_kwargs)
    if "classtype" not in kwargs:
        kwargs["classtype"] = "CEBulk" if "crystalstructure" in kwargs else type(small_bc).__name__
    kwargs["size"] = [int(v) for v in size] if not np.isscalar(size) else int(size)
    large_bc = settings_from_kwargs(kwargs)
    atoms = SimpleAtoms(template_symbols(large_bc))
    CE(atoms, large_bc, eci, initial_cf=init_cf)
    return atoms


def get_atoms_with_ce_calc_JSON(jsonfile, eci=None, size=(1, 1, 1), db_name="temp_db.db"):
    """get_atoms_with_ce_calc with the settings read from a JSON file of
    their arguments and ``classtype`` (ce_calculator.py:104-119)."""
    import json
    from .lattice import settings_from_kwargs
    with open(jsonfile, "r") as infile:
        init_arguments = json.load(infile)
    settings = settings_from_kwargs(init_arguments)
    return get_atoms_with_ce_calc(settings, init_arguments, eci=eci, size=size, db_name=db_name)
