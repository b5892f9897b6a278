"""Move constraints and bias potentials checked on the GPU per trial move:
the reference's cemc/mcmc/mc_constraints.py, cpp/src/pair_constraint.cpp and
cemc/mcmc/bias_potential.py in the forms a device predicate can take.

Each object contributes to the chain's ``cemc_hooks_desc`` (include/cemc_gpu.h)
through ``device_hooks(settings, model)``; ``Montecarlo.add_constraint`` and
``add_bias`` collect them.  On the device a trial move is redrawn while it
violates a constraint (Philox counter (t, k, 0, 0) for proposal k of move t),
up to ``max_constraint_attempts`` proposals (montecarlo.py:970-991), and the
bias of the proposed state joins its energy (montecarlo.py:926-930).

Constraints that are arbitrary Python callables per move cannot run on the
device; they are not accepted.
"""
from __future__ import annotations

import numpy as np


class MCConstraint:
    """Base class (mc_constraints.py:3-22): a constraint is valid for a move
    when it returns True.  Subclasses here also describe themselves for the
    device."""

    name = "GenericConstraint"

    def __call__(self, system_changes):
        return True

    def device_hooks(self, settings, species):
        """Contribution to the chain's hooks: dict with any of
        fixed_species (int mask), allowed (uint8 per site), pair
        ((e1, e2), neighbours array)."""
        raise NotImplementedError


class FixedElement(MCConstraint):
    """No move may change a site from or to ``element`` (mc_constraints.py:84-105)."""

    name = "FixedElement"

    def __init__(self, element=None):
        self.element = element

    def __call__(self, system_changes):
        for change in system_changes:
            if change[1] == self.element or change[2] == self.element:
                return False
        return True

    def device_hooks(self, settings, species):
        if self.element not in species:
            return {}
        return {"fixed_species": 1 << species.index(self.element)}


class ConstrainElementByTag(MCConstraint):
    """Site i may only hold the elements ``element_by_tag[tag_i]``
    (mc_constraints.py:171-190).  Tags come from ``atoms[i].tag`` or the
    ``tags`` argument."""

    name = "ConstrainElementByTag"

    def __init__(self, atoms=None, element_by_tag=(), tags=None):
        self.element_by_tag = [list(e) for e in element_by_tag]
        if tags is None:
            tags = [int(getattr(a, "tag", 0)) for a in atoms]
        self.allowed_elements = [self.element_by_tag[t] for t in tags]

    def __call__(self, system_changes):
        for change in system_changes:
            if change[2] not in self.allowed_elements[change[0]]:
                return False
        return True

    def device_hooks(self, settings, species):
        mask = np.zeros(len(self.allowed_elements), dtype=np.uint8)
        for i, elems in enumerate(self.allowed_elements):
            for e in elems:
                if e in species:
                    mask[i] |= np.uint8(1 << species.index(e))
        return {"allowed": mask}


class FixedSites(MCConstraint):
    """Sites that never change (the device form of FixEdgeLayers,
    mc_constraints.py:108-168, whose geometric test selects such sites)."""

    name = "FixedSites"

    def __init__(self, indices, n_sites):
        self.indices = sorted(set(int(i) for i in indices))
        self.n_sites = int(n_sites)

    def __call__(self, system_changes):
        fixed = set(self.indices)
        return all(change[0] not in fixed for change in system_changes)

    def device_hooks(self, settings, species):
        mask = np.full(self.n_sites, (1 << len(species)) - 1, dtype=np.uint8)
        mask[self.indices] = 0
        return {"allowed": mask}


class FixEdgeLayers(FixedSites):
    """Fix the sites within ``thickness`` of the cell faces
    (mc_constraints.py:108-168): a site is fixed when it lies between one of
    the six plane pairs the reference builds from the cell vectors."""

    name = "FixEdgeLayers"

    def __init__(self, thickness=4.0, atoms=None):
        cell = np.asarray(atoms.get_cell(), dtype=float)
        pos = np.asarray(atoms.get_positions(), dtype=float)
        self.thickness = float(thickness)
        planes = []
        for t, (a, b) in zip((2, 1, 0), ((0, 1), (0, 2), (1, 2))):
            normal = np.cross(cell[a], cell[b])
            unit = cell[t] / np.linalg.norm(cell[t])
            planes.append(((normal, -0.1 * self.thickness * unit), (normal, unit * self.thickness)))
            planes.append(((normal, cell[t] + 0.1 * self.thickness * unit),
                           (normal, cell[t] - self.thickness * unit)))

        def between(x, p1, p2):
            d1 = np.dot(p1[0], x - p1[1]) / np.linalg.norm(p1[0])
            d2 = np.dot(p2[0], x - p2[1]) / np.linalg.norm(p2[0])
            return np.sign(d1 * d2) < 0

        fixed = [i for i in range(len(pos)) if any(between(pos[i], p, q) for p, q in planes)]
        super().__init__(fixed, len(pos))


class PairConstraint(MCConstraint):
    """No site of ``elements[0]`` may have a member of ``elements[1]`` in the
    pair cluster ``cluster_name`` (PairConstraint::elems_in_pair,
    pair_constraint.cpp:14-46).  The reference's Python wrapper passes the
    element list where the cluster name belongs (mc_constraints.py:58-63), so
    there the check finds no cluster and never fires; this class applies the
    C++ check to the named cluster."""

    name = "PairConstraint"

    def __init__(self, calc=None, cluster_name=None, elements=None):
        if calc is None:
            raise ValueError("No calculator object given!")
        if cluster_name is None:
            raise ValueError("No cluster name given!")
        if elements is None:
            raise ValueError("No element list given!")
        if int(cluster_name[1]) != 2:
            raise ValueError("Only pair clusters given.Given cluster has {} elements".format(int(cluster_name[1])))
        if len(elements) != 2:
            raise ValueError("The elements list has to consist of exactly two elements")
        self.calc = calc
        self.cluster_name = cluster_name
        self.elements = list(elements)

    def neighbours(self, settings):
        """(n_sites, n_members) array: the sites of the pair cluster's members
        of each site (trans_matrix(i, member) per site's translation group)."""
        n = settings.n_sites
        groups = settings.index_by_trans_symm
        rows = [None] * n
        for g, sites in enumerate(groups):
            info = settings.cluster_info[g]
            if self.cluster_name not in info:
                raise ValueError("no cluster %s in translation group %d" % (self.cluster_name, g))
            cols = [int(m[0]) for m in info[self.cluster_name]["indices"]]
            for i in sites:
                row = settings.trans_matrix[i]
                rows[i] = [int(row[c]) for c in cols]
        width = {len(r) for r in rows if r is not None}
        if len(width) != 1 or any(r is None for r in rows):
            raise ValueError("every site needs the same number of pair members")
        return np.array(rows, dtype=np.int32)

    def __call__(self, system_changes):
        """Host form of the check, for changes made outside a sampler's
        launch (Montecarlo.insert_symbol): True when, after the changes, no
        site of elements[0] has a member of elements[1] in the cluster
        (the negation of elems_in_pair, pair_constraint.cpp:14-46)."""
        species = list(self.calc.model.species)
        occ = np.array(self.calc.updater.get_occupation(), dtype=np.int64)
        for change in system_changes:
            occ[int(change[0])] = species.index(change[2])
        if getattr(self, "_nb", None) is None:
            self._nb = self.neighbours(self.calc.BC)
        e1, e2 = (species.index(e) for e in self.elements)
        return not bool(np.any((occ[:, None] == e1) & (occ[self._nb] == e2)))

    def device_hooks(self, settings, species):
        e1, e2 = (species.index(e) for e in self.elements)
        return {"pair": ((e1, e2), self.neighbours(settings))}


class BiasPotential:
    """Base class (bias_potential.py:4-28)."""

    def __call__(self, system_changes):
        raise NotImplementedError("Has to be implemented in child classes!")

    def device_table(self, settings, species):
        """(species id, table of n_sites + 1 bias energies per count)."""
        raise NotImplementedError


class ConcentrationBias(BiasPotential):
    """A SampledBiasPotential (bias_potential.py:52-140) whose reaction
    coordinate is the fraction of sites holding ``symbol``: the bias at
    fraction x is the linear interpolation of (reac_crd, free_eng),
    extrapolated beyond the ends as scipy's interp1d(fill_value="extrapolate")
    does.  On the device it is a table over the integer counts."""

    def __init__(self, symbol, reac_crd, free_eng):
        self.symbol = symbol
        self.reac_crd = np.asarray(reac_crd, dtype=float)
        self.free_eng = np.asarray(free_eng, dtype=float)
        if self.reac_crd.ndim != 1 or self.reac_crd.shape != self.free_eng.shape or self.reac_crd.size < 2:
            raise ValueError("reac_crd and free_eng need the same length (at least 2)")

    def get(self, x):
        from scipy.interpolate import interp1d
        f = interp1d(self.reac_crd, self.free_eng, fill_value="extrapolate")
        return np.asarray(f(np.asarray(x, dtype=float)), dtype=float)

    def device_table(self, settings, species):
        n = settings.n_sites
        return species.index(self.symbol), self.get(np.arange(n + 1) / n)


def hooks_for(settings, species, constraints, biases, max_attempts, recycle_waste=False):
    """Merge constraints, biases and the waste-recycling flag into the
    arguments of ChainBatch.set_hooks."""
    out = {"fixed_species": 0, "max_attempts": int(max_attempts), "recycle_waste": bool(recycle_waste)}
    n = settings.n_sites
    allowed = None
    for c in constraints:
        h = c.device_hooks(settings, species)
        out["fixed_species"] |= int(h.get("fixed_species", 0))
        if "allowed" in h:
            a = np.asarray(h["allowed"], dtype=np.uint8)
            if a.shape != (n,):
                raise ValueError("allowed-species mask needs one entry per site")
            allowed = a if allowed is None else (allowed & a)
        if "pair" in h:
            if "pair" in out:
                raise NotImplementedError("one pair constraint per chain on the device")
            out["pair"], out["pair_neighbours"] = h["pair"]
    if allowed is not None:
        out["allowed"] = allowed
    if len(biases) > 1:
        raise NotImplementedError("one bias potential per chain on the device")
    for b in biases:
        out["bias_species"], out["bias_table"] = b.device_table(settings, species)
    return out
