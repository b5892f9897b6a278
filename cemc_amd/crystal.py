"""Cluster-expansion settings for a periodic supercell of any crystal given
by its cell, its sites and its space group: hexagonal close packing, and
crystals of a space group with several Wyckoff positions.

The settings object has the attribute contract ``CEUpdater::init`` reads
(cpp/src/ce_updater.cpp:32-234; cluster_info as parsed by
``Cluster::parse_info_dict``, cpp/src/cluster.cpp:185-232): the same one
``cemc_amd.lattice`` builds for the cubic lattices, here for sites that are
not all related by lattice translations.  In the reference these settings
come from the external ``ase.clease`` package (``CEBulk`` with
crystalstructure="hcp", ``CECrystal`` with a space group); here they are
derived from geometry:

* sites: the repeating unit (its sites given in fractional coordinates)
  repeated L1 x L2 x L3 times;
* the crystal's symmetry operations (R, t) in fractional coordinates, found
  by testing every point operation of the cell's lattice (integer R with
  R^T G R = G) with every translation that takes one site to another; those
  with R = 1 and t != 0 are centering translations;
* translation groups: sites related by lattice or centering translations,
  one group per translation-inequivalent site of the unit (the grouping of
  ``index_by_trans_symm``, ce_updater.cpp:81-97, 733-776);
* clusters containing each group's reference site whose member distances
  are all within the cut-off of their size, grouped into orbits under the
  space group: the orbit key is the smallest, over the operations, of the
  cluster's sorted (unit site, relative cell) list, and the operation that
  attains it fixes a canonical order of the cluster's sites, so a cluster
  contributes consistently from every member (ce_updater.cpp:244-285,
  386-404); the operations that keep the key give ``equiv_sites``;
* trans(i, col): the site that plays col's role when site i is the
  reference site of its group, i.e. col moved by the translation taking the
  group's reference site to i.

Only numpy is used, so the module also runs on the interpreter that drives
the compiled reference when fixtures are written.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

from .lattice import _TransMatrixRows, equivalent_decorations, polynomial_basis

_TOL = 1e-6


def lattice_point_ops(cell):
    """Integer 3x3 matrices R (acting on fractional coordinates) that keep the
    metric G = cell cell^T: the point operations of the lattice."""
    G = np.asarray(cell, dtype=float) @ np.asarray(cell, dtype=float).T
    ops = []
    for v in itertools.product((-1, 0, 1), repeat=9):
        R = np.array(v, dtype=np.int64).reshape(3, 3)
        if abs(round(np.linalg.det(R))) != 1:
            continue
        if np.allclose(R.T @ G @ R, G, atol=1e-8 * max(1.0, np.abs(G).max())):
            ops.append(R)
    return ops


def _frac_key(f):
    """Integer key of fractional coordinates modulo 1."""
    g = np.mod(np.asarray(f, dtype=float), 1.0)
    k = np.rint(g / _TOL).astype(np.int64) % int(round(1.0 / _TOL))
    return tuple(int(x) for x in k)


def space_group_ops(cell, frac, kinds):
    """Operations (R, t), t in [0, 1)^3, that map the sites `frac` of the
    repeating unit onto sites of the same kind, modulo lattice translations."""
    frac = np.asarray(frac, dtype=float)
    lookup = {_frac_key(f): i for i, f in enumerate(frac)}
    out = []
    seen = set()
    for R in lattice_point_ops(cell):
        img0 = R @ frac[0]
        for j in range(len(frac)):
            if kinds[j] != kinds[0]:
                continue
            t = np.mod(frac[j] - img0, 1.0)
            ok = True
            for i, f in enumerate(frac):
                k = lookup.get(_frac_key(R @ f + t))
                if k is None or kinds[k] != kinds[i]:
                    ok = False
                    break
            if not ok:
                continue
            key = (tuple(R.reshape(-1).tolist()), _frac_key(t))
            if key not in seen:
                seen.add(key)
                out.append((R, t))
    return out


def expand_wyckoff(reps, ops):
    """All distinct positions (mod 1) generated from representative positions
    by operations (R, t); returns positions and the index of each one's
    representative."""
    pos, src, seen = [], [], set()
    for r, p in enumerate(reps):
        for R, t in ops:
            q = np.mod(R @ np.asarray(p, dtype=float) + t, 1.0)
            k = _frac_key(q)
            if k in seen:
                continue
            seen.add(k)
            pos.append(q)
            src.append(r)
    return np.array(pos), np.array(src, dtype=np.int64)


class GeneralCrystalSettings:
    """Settings of a periodic supercell of a crystal.

    Parameters
    ----------
    cell : 3x3 array
        Lattice vectors (rows) of the repeating unit.
    frac : (n, 3) array
        Fractional coordinates of the unit's sites.
    basis_elements : list[list[str]]
        Species allowed on each site of the unit (sites with one species are
        background: in ``background_indices``, in no cluster).
    size : int or (int, int, int)
        Repetitions of the unit.
    max_dia_by_size : dict[int, float]
        Largest member distance (length units of `cell`) per cluster size 2-4.
    symmetrize_decorations : bool
        Emit ``equiv_sites`` so symmetry-related decorations share one
        correlation function.
    """

    classtype = "GeneralCrystalSettings"

    def __init__(self, cell, frac, basis_elements, size, max_dia_by_size, symmetrize_decorations=True,
                 kwargs=None):
        if isinstance(size, int):
            size = (size, size, size)
        self.cell = np.asarray(cell, dtype=float)
        self.frac = np.mod(np.asarray(frac, dtype=float), 1.0)
        self.site_elements = [sorted(set(b)) for b in basis_elements]
        if len(self.site_elements) != len(self.frac):
            raise ValueError("one species list per site of the unit is required")
        self.supercell = tuple(int(v) for v in size)
        self.species = sorted(set(e for b in self.site_elements for e in b))
        self.unique_elements = list(self.species)
        self.num_unique_elements = len(self.species)
        self.basis_elements = [list(b) for b in self.site_elements]
        self.symmetrize_decorations = bool(symmetrize_decorations)
        self.self_interaction = False
        self.kwargs = dict(kwargs or {})
        n_unit = len(self.frac)
        kinds = [tuple(b) for b in self.site_elements]
        self.ops = space_group_ops(self.cell, self.frac, kinds)
        self._unit_of = {_frac_key(f): u for u, f in enumerate(self.frac)}
        # centering translations and translation groups of the unit's sites
        cent = [t for R, t in self.ops if np.array_equal(R, np.eye(3, dtype=np.int64))]
        group_of_unit = [-1] * n_unit
        ref_units = []
        for u in range(n_unit):
            if group_of_unit[u] >= 0:
                continue
            g = len(ref_units)
            ref_units.append(u)
            for t in cent:
                v = self._unit_of[_frac_key(self.frac[u] + t)]
                if group_of_unit[v] < 0:
                    group_of_unit[v] = g
        self.active_unit = [len(b) > 1 for b in self.site_elements]
        self.unit_group = group_of_unit
        L1, L2, L3 = self.supercell
        n_cells = L1 * L2 * L3
        self.n_unit = n_unit
        self.n_sites = n_cells * n_unit
        c = np.arange(n_cells)
        self.cell_of_site = np.repeat(np.stack([c // (L2 * L3), (c // L3) % L2, c % L3], axis=1), n_unit, axis=0)
        self.unit_of_site = np.tile(np.arange(n_unit), n_cells)
        self.sublattice_of_site = self.unit_of_site  # basis_elements is per unit site
        self.background_indices = [int(i) for i in range(self.n_sites) if not self.active_unit[i % n_unit]]
        # groups in order of their reference unit site; background units form
        # no group of their own
        self.group_refs = [u for u in ref_units if self.active_unit[u]]
        gidx = {u: k for k, u in enumerate(self.group_refs)}
        self.index_by_trans_symm = [[] for _ in self.group_refs]
        for i in range(self.n_sites):
            u = i % n_unit
            if self.active_unit[u]:
                self.index_by_trans_symm[gidx[ref_units[group_of_unit[u]]]].append(i)
        phi = polynomial_basis(len(self.species))
        self.basis = phi
        self.basis_functions = [{sp: float(phi[d, k]) for k, sp in enumerate(self.species)}
                                for d in range(phi.shape[0])]
        self._build({int(k): float(v) for k, v in dict(max_dia_by_size).items()})

    # ----------------------------------------------------------- geometry
    def _cart(self, cells, units):
        return (np.asarray(cells, dtype=float) + self.frac[np.asarray(units)]) @ self.cell

    def _decompose(self, f):
        """(unit site, integer cell) of fractional position f."""
        cellv = np.floor(f + _TOL)
        u = self._unit_of.get(_frac_key(f - cellv))
        if u is None:
            raise ValueError("position is not a site of the crystal")
        cellv = np.rint(f - self.frac[u]).astype(np.int64)
        return u, cellv

    def site_of(self, unit, cellv):
        L1, L2, L3 = self.supercell
        ci = ((int(cellv[0]) % L1) * L2 + (int(cellv[1]) % L2)) * L3 + (int(cellv[2]) % L3)
        return ci * self.n_unit + int(unit)

    def _neighbours(self, u0, rmax):
        """(unit, cell) of active sites within rmax of unit site u0 in cell 0."""
        p0 = self._cart([0, 0, 0], [u0])[0]
        inv = np.linalg.inv(self.cell)
        span = np.ceil(rmax * np.linalg.norm(inv, axis=0)).astype(int) + 1
        out = []
        for n in itertools.product(*[range(-s, s + 1) for s in span]):
            for u in range(self.n_unit):
                if not self.active_unit[u] or (u == u0 and n == (0, 0, 0)):
                    continue
                d = np.linalg.norm(self._cart(n, [u])[0] - p0)
                if d <= rmax + 1e-9:
                    out.append((u, np.array(n, dtype=np.int64), d))
        return out

    def _key(self, pts):
        """Sorted (unit, relative cell) list of points given as (unit, cell)."""
        srt = sorted(range(len(pts)), key=lambda k: (pts[k][0], tuple(pts[k][1])))
        base = pts[srt[0]][1]
        return tuple((pts[k][0],) + tuple(int(x) for x in pts[k][1] - base) for k in srt), srt

    def _build(self, cut):
        info = [dict() for _ in self.group_refs]
        for g, u in enumerate(self.group_refs):
            info[g]["c0"] = {"name": "c0", "descriptor": "empty", "size": 0, "ref_indx": int(u),
                             "symm_group": g, "max_cluster_dia": 0.0, "indices": [], "order": [],
                             "equiv_sites": []}
            info[g]["c1"] = {"name": "c1", "descriptor": "point_cluster", "size": 1, "ref_indx": int(u),
                             "symm_group": g, "max_cluster_dia": 0.0, "indices": [], "order": [[0]],
                             "equiv_sites": []}
        found = []  # (size, dia, orbit key, group, [(sites of cluster, canonical ranks)], equiv)
        for size in sorted(cut):
            if size not in (2, 3, 4):
                raise ValueError("Only cluster sizes 2, 3 and 4 are supported")
            rmax = cut[size]
            for g, u0 in enumerate(self.group_refs):
                nb = self._neighbours(u0, rmax)
                carts = [self._cart(n, [u])[0] for u, n, _ in nb]
                orbits = {}
                for cmb in itertools.combinations(range(len(nb)), size - 1):
                    if any(np.linalg.norm(carts[a] - carts[b]) > rmax + 1e-9
                           for a, b in itertools.combinations(cmb, 2)):
                        continue
                    pts = [(u0, np.zeros(3, dtype=np.int64))] + [(nb[k][0], nb[k][1]) for k in cmb]
                    best, ranks_best, stab = None, None, []
                    for R, t in self.ops:
                        img = [self._decompose(R @ (n + self.frac[u]) + t) for u, n in pts]
                        key, srt = self._key(img)
                        ranks = [0] * size
                        for pos, k in enumerate(srt):
                            ranks[k] = pos
                        if best is None or key < best:
                            best, ranks_best, stab = key, ranks, [ranks]
                        elif key == best:
                            stab.append(ranks)
                    cart = [self._cart(n, [u])[0] for u, n in pts]
                    dia = max(np.linalg.norm(a - b) for a, b in itertools.combinations(cart, 2))
                    ent = orbits.setdefault(best, {"dia": dia, "subs": [], "stab": stab, "ranks": ranks_best})
                    ent["subs"].append((pts, ranks_best))
                for key, ent in orbits.items():
                    found.append((size, round(ent["dia"], 6), key, g, ent))
        # names: per size by diameter, then orbit key
        keys = sorted({(f[0], f[1], f[2]) for f in found})
        names = {}
        for size in sorted({k[0] for k in keys}):
            same = [k for k in keys if k[0] == size]
            dias = sorted({k[1] for k in same})
            for di, d in enumerate(dias):
                for k_i, k in enumerate([k for k in same if k[1] == d]):
                    names[k] = "c%d_d%04d_%d" % (size, di, k_i)
        self.orbit_names = names
        self.orbit_diameter = {names[k]: k[1] for k in keys}
        self.orbit_sizes = {names[k]: k[0] for k in keys}
        self.orbit_equiv = {}
        used = set()
        for size, dia, key, g, ent in found:
            name = names[(size, dia, key)]
            members, order = [], []
            for pts, ranks in ent["subs"]:
                by_pos = sorted(range(size), key=lambda k: ranks[k])
                r = ranks[0]
                mem = [self.site_of(pts[k][0], pts[k][1]) for k in by_pos if k != 0]
                members.append(mem)
                order.append([k + 1 if k < r else (0 if k == r else k) for k in range(size)])
                used.update(mem)
            ref_site = int(self.group_refs[g])
            if any(ref_site in mem or len(set(mem)) != len(mem) for mem in members):
                self.self_interaction = True
            eqs = self._equiv_groups(ent["ranks"], ent["stab"]) if self.symmetrize_decorations else []
            self.orbit_equiv[name] = eqs
            info[g][name] = {"name": name, "descriptor": "%d-body d=%.6f" % (size, dia), "size": size,
                             "ref_indx": ref_site, "symm_group": g, "max_cluster_dia": float(ent["dia"]),
                             "indices": members, "order": order, "equiv_sites": eqs}
        self.cluster_info = info
        # trans(i, col): col moved by the translation taking i's group
        # reference site to i
        cols = np.array(sorted(used), dtype=np.int64)
        self.trans_cols = cols.astype(np.int32)
        col_pos = np.array([self.cell_of_site[c] + self.frac[self.unit_of_site[c]] for c in cols])
        tab = np.zeros((self.n_sites, len(cols)), dtype=np.int32)
        for i in range(self.n_sites):
            u = int(self.unit_of_site[i])
            if not self.active_unit[u]:
                tab[i] = -1
                continue
            g_ref = self._ref_unit_of(u)
            shift = self.cell_of_site[i] + self.frac[u] - self.frac[g_ref]
            for j, p in enumerate(col_pos):
                uu, cc = self._decompose(p + shift)
                tab[i, j] = self.site_of(uu, cc)
        self.trans_table = np.ascontiguousarray(tab)
        self.trans_matrix = _TransMatrixRows(self.trans_cols, self.trans_table)

    def _ref_unit_of(self, u):
        for r in self.group_refs:
            if self.unit_group[r] == self.unit_group[u]:
                return r
        raise ValueError("unit site %d has no group" % u)

    @staticmethod
    def _equiv_groups(base, stab):
        """Canonical positions equivalent under the cluster's stabiliser."""
        n = len(base)
        parent = list(range(n))

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        for other in stab:
            for site in range(n):
                a, b = find(int(base[site])), find(int(other[site]))
                if a != b:
                    parent[a] = b
        groups = {}
        for pos in range(n):
            groups.setdefault(find(pos), []).append(pos)
        return [sorted(gr) for gr in sorted(groups.values()) if len(gr) > 1]

    # ---------------------------------------------------------- CF naming
    def decorations(self, name):
        nb = len(self.basis_functions)
        size = self.orbit_sizes[name]
        decos = ["".join(str(v) for v in d) for d in itertools.product(range(nb), repeat=size)]
        eqs = self.orbit_equiv.get(name, [])
        if not self.symmetrize_decorations or not eqs:
            return decos
        keep, seen = [], set()
        for d in decos:
            cls = tuple(sorted("".join(str(v) for v in e)
                               for e in equivalent_decorations([int(c) for c in d], eqs)))
            if cls in seen:
                continue
            seen.add(cls)
            keep.append(d)
        return keep

    def cf_names(self):
        names = ["c0"] + ["c1_%d" % d for d in range(len(self.basis_functions))]
        for name in sorted(self.orbit_sizes):
            names += ["%s_%s" % (name, d) for d in self.decorations(name)]
        return sorted(names)

    def cf_diameter(self, name):
        if name.startswith("c0") or name.startswith("c1"):
            return 0.0
        return self.orbit_diameter[name.rsplit("_", 1)[0]]


def hcp_cell(a=1.0, c_over_a=math.sqrt(8.0 / 3.0)):
    """Hexagonal cell and the two HCP sites (P6_3/mmc, Wyckoff 2c)."""
    cell = np.array([[a, 0.0, 0.0], [-0.5 * a, 0.5 * math.sqrt(3.0) * a, 0.0], [0.0, 0.0, c_over_a * a]])
    frac = np.array([[1.0 / 3.0, 2.0 / 3.0, 0.25], [2.0 / 3.0, 1.0 / 3.0, 0.75]])
    return cell, frac


class HCPSettings(GeneralCrystalSettings):
    """Settings of a periodic HCP supercell (CLEASE ``CEBulk`` with
    crystalstructure="hcp"): two sites per hexagonal cell, one translation
    group each (they are related by the 6_3 screw and mirrors, not by a
    translation), clusters up to `max_dia_by_size` (units of a)."""

    classtype = "HCPSettings"

    def __init__(self, size, species, max_dia_by_size, c_over_a=math.sqrt(8.0 / 3.0), symmetrize_decorations=True):
        cell, frac = hcp_cell(1.0, c_over_a)
        kw = {"size": list(size) if not isinstance(size, int) else [size] * 3, "species": sorted(species),
              "max_dia_by_size": {str(k): float(v) for k, v in dict(max_dia_by_size).items()},
              "c_over_a": float(c_over_a), "symmetrize_decorations": bool(symmetrize_decorations)}
        super().__init__(cell, frac, [sorted(species)] * 2, size, max_dia_by_size, symmetrize_decorations, kw)


def tetrahedral_point_ops():
    """The 24 operations of -43m (T_d) as 3x3 integer matrices in cubic axes:
    the signed permutations that keep the tetrahedron (1,1,1), (1,-1,-1),
    (-1,1,-1), (-1,-1,1)."""
    tet = {(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)}
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=np.int64)
            for row, (col, s) in enumerate(zip(perm, signs)):
                m[row, col] = s
            if {tuple(int(x) for x in m @ np.array(v)) for v in tet} == tet:
                ops.append(m)
    return ops


class Spacegroup217Settings(GeneralCrystalSettings):
    """A body-centred cubic crystal of space group I-43m (217) from Wyckoff
    representatives (CLEASE ``CECrystal(spacegroup=217, basis=...)``):
    positions generated by the 24 operations of -43m and the centring
    (1/2, 1/2, 1/2) in the conventional cubic cell, which repeats
    L1 x L2 x L3 times.  Sites related by the centring share a translation
    group."""

    classtype = "Spacegroup217Settings"

    def __init__(self, size, a, wyckoff, wyckoff_elements, max_dia_by_size, symmetrize_decorations=True):
        ops = [(R, np.zeros(3)) for R in tetrahedral_point_ops()]
        ops += [(R, np.full(3, 0.5)) for R in tetrahedral_point_ops()]
        frac, src = expand_wyckoff(wyckoff, ops)
        order = np.lexsort((frac[:, 2], frac[:, 1], frac[:, 0], src))
        frac, src = frac[order], src[order]
        elements = [sorted(wyckoff_elements[int(r)]) for r in src]
        kw = {"size": list(size) if not isinstance(size, int) else [size] * 3, "a": float(a),
              "wyckoff": [list(map(float, w)) for w in wyckoff], "wyckoff_elements": [list(e) for e in wyckoff_elements],
              "max_dia_by_size": {str(k): float(v) for k, v in dict(max_dia_by_size).items()},
              "symmetrize_decorations": bool(symmetrize_decorations)}
        super().__init__(np.eye(3) * float(a), frac, elements, size, max_dia_by_size, symmetrize_decorations, kw)
