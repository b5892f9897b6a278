"""Cluster-expansion settings for periodic supercells of the cubic Bravais
lattices (FCC, BCC, SC).

This module builds, from lattice geometry alone, the settings object that the
reference updater consumes at initialisation (the attribute contract read by
``CEUpdater::init``, cpp/src/ce_updater.cpp:32-234, and by
``Cluster::parse_info_dict``, cpp/src/cluster.cpp:185-232):

* ``unique_elements``        list[str]            (ce_updater.cpp:70-75)
* ``index_by_trans_symm``    list[list[int]]      (ce_updater.cpp:81-97, 733-776)
* ``background_indices``     list[int]            (ce_updater.cpp:92-94, 1015-1027)
* ``num_unique_elements``    int                  (ce_updater.cpp:106-113)
* ``cluster_info``           list[dict[name -> info]] with ``ref_indx, size,
  max_cluster_dia, symm_group, name, descriptor, indices, order, equiv_sites``
* ``basis_functions``        list[dict[symbol -> float]] (ce_updater.cpp:157-181)
* ``trans_matrix``           list[dict[col -> site]]  (ce_updater.cpp:899-979)

In the reference these come from the external ``ase.clease`` package, which is
not part of the reference tree.  Here they are derived directly:

* sites of an ``L1 x L2 x L3`` supercell of the primitive cell of the FCC,
  BCC or SC lattice, with integer coordinates in units of a/2 (a = cubic
  lattice constant);
* clusters containing the reference site, grouped into orbits under the 48
  operations of the cubic point group plus lattice translations;
* a canonical site order per cluster that does not depend on which member is
  the reference site, so each cluster contributes consistently from all of its
  members (the property the incremental update relies on,
  ce_updater.cpp:244-285 and 386-404);
* symmetry-equivalent positions inside each cluster (``equiv_sites``);
* an orthonormal polynomial site basis.

The module uses only numpy so it also runs on the older interpreter used to
drive the compiled reference when test fixtures are generated.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

# Primitive vectors (rows) of the cubic Bravais lattices in units of a/2.
FCC_PRIM = np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.int64)
BCC_PRIM = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], dtype=np.int64)
SC_PRIM = np.array([[2, 0, 0], [0, 2, 0], [0, 0, 2]], dtype=np.int64)
PRIMITIVE = {"fcc": FCC_PRIM, "bcc": BCC_PRIM, "sc": SC_PRIM}

# Boltzmann constant in eV/K (CODATA 2014, the value ase.units.kB carried when
# the reference was written).
KB_EV = 8.617330337217213e-05


def cubic_point_group():
    """The 48 signed 3x3 permutation matrices of the cubic group Oh."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=np.int64)
            for row, (col, s) in enumerate(zip(perm, signs)):
                m[row, col] = s
            ops.append(m)
    return np.array(ops)


_OPS = cubic_point_group()
_CODE_OFF = 64
_CODE_BASE = 128


def _encode(points):
    """Monotone (lexicographic) integer code of integer 3-vectors (..., 3)."""
    p = points + _CODE_OFF
    return (p[..., 0] * _CODE_BASE + p[..., 1]) * _CODE_BASE + p[..., 2]


_CODE_ZERO = _encode(np.zeros(3, dtype=np.int64))


def _on_lattice(pts, lattice):
    """Mask of integer points (a/2 units) that are sites of the lattice: FCC
    even coordinate sum, BCC coordinates of one parity, SC even coordinates."""
    if lattice == "fcc":
        return (pts.sum(axis=1) % 2) == 0
    if lattice == "bcc":
        return ((pts[:, 0] - pts[:, 1]) % 2 == 0) & ((pts[:, 1] - pts[:, 2]) % 2 == 0)
    if lattice == "sc":
        return np.all(pts % 2 == 0, axis=1)
    raise ValueError("unknown lattice %r (fcc, bcc or sc)" % (lattice,))


def lattice_points_within(d2_max, lattice="fcc"):
    """Lattice points (a/2 units) with 0 < |p|^2 <= d2_max, by (|p|^2, x, y, z)."""
    r = int(math.isqrt(int(d2_max))) + 1
    g = np.arange(-r, r + 1)
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    pts = pts[_on_lattice(pts, lattice)]
    d2 = (pts * pts).sum(axis=1)
    keep = (d2 > 0) & (d2 <= d2_max)
    pts = pts[keep]
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], (pts * pts).sum(axis=1)))
    return pts[order]


def fcc_points_within(d2_max):
    """FCC lattice points (a/2 units) with 0 < |p|^2 <= d2_max."""
    return lattice_points_within(d2_max, "fcc")


def lattice_shells(n_shells, lattice="fcc"):
    """Squared radii (a/2 units) of the first ``n_shells`` neighbour shells."""
    pts = lattice_points_within(12 * n_shells + 12, lattice)
    d2 = np.unique((pts * pts).sum(axis=1))
    return [int(v) for v in d2[:n_shells]]


def fcc_shells(n_shells):
    """Squared radii (a/2 units) of the first ``n_shells`` FCC neighbour shells."""
    return lattice_shells(n_shells, "fcc")


def _clusters_with_origin(size, d2_cut, lattice="fcc"):
    """All clusters of ``size`` sites containing the origin whose pairwise squared
    distances are all <= d2_cut.  Returns an int array (n_clusters, size, 3) with
    the origin first."""
    cand = lattice_points_within(d2_cut, lattice)
    n_c = len(cand)
    diff = cand[:, None, :] - cand[None, :, :]
    ok = (diff * diff).sum(axis=2) <= d2_cut
    out = []
    for combo in itertools.combinations(range(n_c), size - 1):
        good = True
        for a, b in itertools.combinations(combo, 2):
            if not ok[a, b]:
                good = False
                break
        if good:
            out.append(combo)
    if not out:
        return np.zeros((0, size, 3), dtype=np.int64)
    idx = np.array(out, dtype=np.int64)
    pts = cand[idx]
    origin = np.zeros((len(idx), 1, 3), dtype=np.int64)
    return np.concatenate([origin, pts], axis=1)


def _canonical(clusters):
    """Translation/rotation-invariant orbit keys and canonical site ranks.

    For every cluster and every point operation R the rotated points are sorted
    lexicographically and translated so the smallest point sits at the origin;
    the key is the smallest such sorted tuple over R.  The first R (in the fixed
    order of ``_OPS``) that attains the key fixes the canonical position of each
    site: its rank in the sorted rotated tuple.  Because lexicographic order is
    preserved by translation, the same physical cluster gets the same canonical
    order whichever member is placed at the origin.
    """
    n_cl, n, _ = clusters.shape
    rot = np.einsum("rij,cnj->crni", _OPS, clusters)  # (n_cl, 48, n, 3)
    codes = _encode(rot)  # (n_cl, 48, n)
    ranks = np.argsort(np.argsort(codes, axis=2, kind="stable"), axis=2, kind="stable")
    srt = np.sort(codes, axis=2)
    rel = srt - srt[..., :1] + _CODE_ZERO  # translation to the smallest point
    packed = np.zeros((n_cl, len(_OPS)), dtype=np.int64)
    shift = 0
    for j in range(1, n):
        packed += rel[..., j].astype(np.int64) << shift
        shift += 21
    key = packed
    best = key.min(axis=1)
    first_op = np.argmax(key == best[:, None], axis=1)
    canon_rank = ranks[np.arange(n_cl), first_op]  # (n_cl, n)
    stab = key == best[:, None]
    return best, canon_rank, stab, ranks


def _equiv_groups(ranks_all, stab_row, first_op):
    """Groups of equivalent canonical positions from the stabiliser of one cluster."""
    base = ranks_all[first_op]
    n = len(base)
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for r in np.nonzero(stab_row)[0]:
        other = ranks_all[r]
        for site in range(n):
            a, b = find(int(base[site])), find(int(other[site]))
            if a != b:
                parent[a] = b
    groups = {}
    for pos in range(n):
        groups.setdefault(find(pos), []).append(pos)
    return [sorted(g) for g in sorted(groups.values()) if len(g) > 1]


class Orbit:
    """One symmetry class of clusters as seen from the reference site."""

    def __init__(self, size, d2, key, offsets, positions, equiv_sites):
        self.size = size
        self.d2 = d2  # max squared intra-cluster distance, (a/2)^2 units
        self.key = key
        # offsets[m] = (size, 3) integer positions of sub-cluster m, origin = ref
        self.offsets = offsets
        # positions[m][s] = canonical position of site s of offsets[m]
        self.positions = positions
        self.equiv_sites = equiv_sites
        self.name = None

    @property
    def diameter(self):
        """Maximum intra-cluster distance in units of the cubic lattice constant."""
        return math.sqrt(self.d2) / 2.0

    @property
    def multiplicity(self):
        return len(self.offsets)


def enumerate_orbits(max_shell_by_size, lattice="fcc"):
    """Cluster orbits of a cubic Bravais lattice (default FCC).

    ``max_shell_by_size`` maps cluster size (2, 3 or 4) to the outermost
    neighbour shell allowed for every pair of sites in the cluster.
    Orbits are returned sorted by (size, diameter, key) and named
    ``c{size}_d{diameter index:04d}_{k}``.
    """
    if not max_shell_by_size:
        return []
    n_sh = max(max_shell_by_size.values())
    shells = lattice_shells(n_sh, lattice)
    orbits = []
    for size in sorted(max_shell_by_size):
        if size not in (2, 3, 4):
            raise ValueError("Only cluster sizes 2, 3 and 4 are supported")
        d2_cut = shells[max_shell_by_size[size] - 1]
        cl = _clusters_with_origin(size, d2_cut, lattice)
        if len(cl) == 0:
            continue
        best, canon_rank, stab, ranks = _canonical(cl)
        first_op = np.argmax(stab, axis=1)
        d = cl[:, :, None, :] - cl[:, None, :, :]
        d2max = (d * d).sum(axis=3).reshape(len(cl), -1).max(axis=1)
        for key in np.unique(best):
            sel = np.nonzero(best == key)[0]
            rep = sel[0]
            eq = _equiv_groups(ranks[rep], stab[rep], first_op[rep])
            orbits.append(Orbit(size, int(d2max[rep]), int(key), cl[sel],
                                canon_rank[sel], eq))
    orbits.sort(key=lambda o: (o.size, o.d2, o.key))
    for size in sorted({o.size for o in orbits}):
        same = [o for o in orbits if o.size == size]
        dias = sorted({o.d2 for o in same})
        for di, d2 in enumerate(dias):
            for k, o in enumerate([o for o in same if o.d2 == d2]):
                o.name = "c%d_d%04d_%d" % (size, di, k)
    return orbits


def species_spins(n_species):
    """Integer spin values assigned to the species in sorted-name order."""
    if n_species % 2 == 0:
        half = n_species // 2
        pos = list(range(half, 0, -1))
        return pos + [-v for v in reversed(pos)]
    half = n_species // 2
    return list(range(half, -half - 1, -1))


def polynomial_basis(n_species):
    """Orthonormal polynomial site basis.

    Returns an array ``phi[d, s]`` (d < n_species-1) such that
    (1/n_s) sum_s phi[i, s] phi[j, s] = delta_ij and sum_s phi[d, s] = 0.
    For two species this is phi = (+1, -1).
    """
    spins = np.array(species_spins(n_species), dtype=np.float64)
    vecs = [np.ones(n_species)]
    out = []
    for power in range(1, n_species):
        v = spins ** power
        for u in vecs:
            v = v - (np.dot(v, u) / np.dot(u, u)) * u
        v = v / math.sqrt(np.dot(v, v) / n_species)
        if v[0] < 0:
            v = -v
        vecs.append(v)
        out.append(v)
    return np.array(out)


def equivalent_decorations(deco, equiv_sites):
    """All decorations obtained by permuting decoration digits inside each group
    of equivalent positions (Cartesian product over groups, duplicates kept)."""
    result = [list(deco)]
    for group in equiv_sites:
        nxt = []
        for d in result:
            vals = [d[p] for p in group]
            for perm in itertools.permutations(vals):
                e = list(d)
                for p, v in zip(group, perm):
                    e[p] = v
                nxt.append(e)
        result = nxt
    return result


class _TransMatrixRows:
    """Read-only list-of-dict view of the translation matrix (only used columns)."""

    def __init__(self, cols, table):
        self._cols = [int(c) for c in cols]
        self._table = table

    def __len__(self):
        return self._table.shape[0]

    def __getitem__(self, i):
        row = self._table[i]
        return {c: int(v) for c, v in zip(self._cols, row)}

    def to_list(self):
        return [self[i] for i in range(len(self))]


class CubicSettings:
    """Settings of a periodic supercell cluster expansion on a cubic Bravais
    lattice.

    Parameters
    ----------
    size : int or (int, int, int)
        Supercell repetitions of the primitive cell.
    species : list[str]
        Chemical species; ids follow sorted-name order as in the reference
        (cpp/src/symbols_with_numbers.cpp:11-21).
    max_shell_by_size : dict[int, int]
        Outermost neighbour shell per cluster size, e.g. ``{2: 4, 3: 2}``.
    symmetrize_decorations : bool
        Emit ``equiv_sites`` so decorations related by cluster symmetry share
        one correlation function.  With ``False`` every decoration is its own
        correlation function and ``equiv_sites`` is empty.
    lattice : str
        "fcc", "bcc" or "sc".
    """

    classtype = "CubicSettings"

    def __init__(self, size, species, max_shell_by_size, symmetrize_decorations=True, lattice="fcc"):
        if isinstance(size, int):
            size = (size, size, size)
        lattice = str(lattice).lower()
        if lattice not in PRIMITIVE:
            raise ValueError("unknown lattice %r (fcc, bcc or sc)" % (lattice,))
        self.lattice = lattice
        self.prim = PRIMITIVE[lattice]
        # integer inverse of the primitive matrix: cell = point @ adj / det
        self._det = int(round(np.linalg.det(self.prim)))
        self._adj = np.rint(np.linalg.inv(self.prim) * self._det).astype(np.int64)
        self.dims = tuple(int(v) for v in size)
        self.species = sorted(species)
        self.unique_elements = list(self.species)
        self.num_unique_elements = len(self.species)
        self.basis_elements = [list(self.species)]
        self.n_sites = int(np.prod(self.dims))
        self.background_indices = []
        self.index_by_trans_symm = [list(range(self.n_sites))]
        self.symmetrize_decorations = bool(symmetrize_decorations)
        self.self_interaction = False
        # arguments that rebuild these settings (CE.save / CE.load)
        self.kwargs = {"size": list(self.dims), "species": list(self.species),
                       "max_shell_by_size": {str(k): int(v) for k, v in dict(max_shell_by_size).items()},
                       "symmetrize_decorations": self.symmetrize_decorations, "lattice": self.lattice}
        self.orbits = enumerate_orbits(dict(max_shell_by_size), self.lattice)
        phi = polynomial_basis(len(self.species))
        self.basis = phi
        self.basis_functions = [
            {s: float(phi[d, k]) for k, s in enumerate(self.species)}
            for d in range(phi.shape[0])
        ]
        self._build_sites()
        self._build_cluster_info()

    # ------------------------------------------------------------------ sites
    def site_index(self, cells):
        """Site index of integer cell coordinates (..., 3), wrapped periodically."""
        L1, L2, L3 = self.dims
        c = np.asarray(cells, dtype=np.int64)
        return ((c[..., 0] % L1) * L2 + (c[..., 1] % L2)) * L3 + (c[..., 2] % L3)

    def _build_sites(self):
        L1, L2, L3 = self.dims
        n = np.arange(self.n_sites)
        self.cells = np.stack([n // (L2 * L3), (n // L3) % L2, n % L3], axis=1)
        self.positions = self.cells @ self.prim  # a/2 units

    def _to_cell(self, points):
        """Integer cell coordinates of lattice points (a/2 units)."""
        num = np.asarray(points, dtype=np.int64) @ self._adj
        if np.any(num % self._det):
            raise ValueError("point is not a lattice site")
        return num // self._det

    # -------------------------------------------------------- cluster info
    def _build_cluster_info(self):
        info = {}
        info["c0"] = {"name": "c0", "descriptor": "empty", "size": 0, "ref_indx": 0,
                      "symm_group": 0, "max_cluster_dia": 0.0, "indices": [],
                      "order": [], "equiv_sites": []}
        info["c1"] = {"name": "c1", "descriptor": "point_cluster", "size": 1,
                      "ref_indx": 0, "symm_group": 0, "max_cluster_dia": 0.0,
                      "indices": [], "order": [[0]], "equiv_sites": []}
        used_cols = set()
        self.orbit_members = {}
        for orb in self.orbits:
            n = orb.size
            members = []
            order = []
            for m in range(orb.multiplicity):
                off = orb.offsets[m]
                pos = orb.positions[m]
                # members listed in canonical order, reference site skipped
                by_pos = np.argsort(pos)
                r = int(pos[0])
                mem_sites = [int(self.site_index(self._to_cell(off[s]))) for s in by_pos if s != 0]
                members.append(mem_sites)
                order.append([k + 1 if k < r else (0 if k == r else k) for k in range(n)])
                used_cols.update(mem_sites)
            if any(0 in mem or len(set(mem)) != len(mem) for mem in members):
                self.self_interaction = True
            eq = orb.equiv_sites if self.symmetrize_decorations else []
            info[orb.name] = {"name": orb.name, "descriptor": "%d-body d2=%d" % (n, orb.d2),
                              "size": n, "ref_indx": 0, "symm_group": 0,
                              "max_cluster_dia": orb.diameter, "indices": members,
                              "order": order, "equiv_sites": [list(g) for g in eq]}
            self.orbit_members[orb.name] = members
        self.cluster_info = [info]
        cols = np.array(sorted(used_cols), dtype=np.int64)
        # trans(i, col): the site playing the role of `col` when site i is the
        # reference site (translation of the supercell taking site 0 to site i)
        col_cells = self.cells[cols]
        self.trans_cols = cols.astype(np.int32)
        self.trans_col_cells = col_cells.astype(np.int32)
        tab = self.site_index(self.cells[:, None, :] + col_cells[None, :, :])
        self.trans_table = np.ascontiguousarray(tab.astype(np.int32))
        self.trans_matrix = _TransMatrixRows(self.trans_cols, self.trans_table)

    # ------------------------------------------------------- CF naming
    def decorations(self, orbit):
        """Decoration strings that carry their own correlation function."""
        nb = len(self.basis_functions)
        decos = ["".join(str(v) for v in d) for d in itertools.product(range(nb), repeat=orbit.size)]
        if not self.symmetrize_decorations or not orbit.equiv_sites:
            return decos
        keep = []
        seen = set()
        for d in decos:
            cls = tuple(sorted("".join(str(v) for v in e)
                               for e in equivalent_decorations([int(c) for c in d], orbit.equiv_sites)))
            if cls in seen:
                continue
            seen.add(cls)
            keep.append(d)
        return keep

    def cf_names(self):
        """All correlation-function names, sorted as the reference sorts them."""
        names = ["c0"] + ["c1_%d" % d for d in range(len(self.basis_functions))]
        for orb in self.orbits:
            names += ["%s_%s" % (orb.name, d) for d in self.decorations(orb)]
        return sorted(names)

    def cf_diameter(self, name):
        """Cluster diameter (units of a) of a correlation function name (0 for c0/c1)."""
        if name.startswith("c0") or name.startswith("c1"):
            return 0.0
        prefix = name.rsplit("_", 1)[0]
        for orb in self.orbits:
            if orb.name == prefix:
                return orb.diameter
        raise KeyError(name)

    def periodic_descriptor(self):
        """Closed-form description of the translation matrix: supercell dims and,
        per used column, its cell offset.  trans(i, col) = site(cell(i) + off(col))."""
        return {"dims": self.dims, "cols": self.trans_cols, "col_cells": self.trans_col_cells}


# Two-sublattice cubic crystals: primitive vectors and basis (a/2 units).
# Rocksalt (B1): FCC lattice, basis (0,0,0) and (1,0,0); the sublattice of a
# point is the parity of its coordinate sum.  CsCl (B2): simple-cubic
# lattice, basis (0,0,0) and (1,1,1); the sublattice is the parity of x.
CRYSTALS = {
    "rocksalt": (FCC_PRIM, np.array([[0, 0, 0], [1, 0, 0]], dtype=np.int64), "sc_half"),
    "cscl": (SC_PRIM, np.array([[0, 0, 0], [1, 1, 1]], dtype=np.int64), "bcc"),
}


def _sublattice(points, structure):
    p = np.asarray(points, dtype=np.int64)
    if structure == "rocksalt":
        return p.sum(axis=-1) % 2
    return p[..., 0] % 2


def _crystal_points_within(d2_max, structure):
    """Points of the combined (all-sublattice) point set with 0 < |p|^2 <=
    d2_max: every integer point for rocksalt, BCC points for CsCl."""
    r = int(math.isqrt(int(d2_max))) + 1
    g = np.arange(-r, r + 1)
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    if structure == "cscl":
        pts = pts[_on_lattice(pts, "bcc")]
    d2 = (pts * pts).sum(axis=1)
    pts = pts[(d2 > 0) & (d2 <= d2_max)]
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], (pts * pts).sum(axis=1)))
    return pts[order]


class CrystalSettings:
    """Settings of a periodic two-sublattice cubic crystal (CLEASE
    ``CECrystal`` of rocksalt or CsCl type) with one translation group per
    active sublattice.

    Parameters
    ----------
    size : int or (int, int, int)
        Supercell repetitions of the primitive cell (two sites per cell).
    basis_elements : [list[str], list[str]]
        Species allowed on each sublattice.  A sublattice with one species is
        background: its sites are in ``background_indices``, in no
        translation group and in no cluster (ce_updater.cpp:81-97, 733-776,
        1015-1027).
    max_shell_by_size : dict[int, int]
        Outermost neighbour shell (of the combined point set) per cluster size.
    structure : "rocksalt" or "cscl"
    """

    classtype = "CrystalSettings"

    def __init__(self, size, basis_elements, max_shell_by_size, structure="rocksalt",
                 symmetrize_decorations=True):
        if isinstance(size, int):
            size = (size, size, size)
        structure = str(structure).lower()
        if structure not in CRYSTALS:
            raise ValueError("unknown structure %r (rocksalt or cscl)" % (structure,))
        if len(basis_elements) != 2:
            raise ValueError("two sublattices are required")
        self.structure = structure
        self.prim, self.basis_pos, _ = CRYSTALS[structure]
        self._det = int(round(np.linalg.det(self.prim)))
        self._adj = np.rint(np.linalg.inv(self.prim) * self._det).astype(np.int64)
        self.supercell = tuple(int(v) for v in size)
        self.basis_elements = [sorted(set(b)) for b in basis_elements]
        self.species = sorted(set(e for b in self.basis_elements for e in b))
        self.unique_elements = list(self.species)
        self.num_unique_elements = len(self.species)
        self.symmetrize_decorations = bool(symmetrize_decorations)
        self.self_interaction = False
        self.kwargs = {"size": list(self.supercell), "basis_elements": [list(b) for b in self.basis_elements],
                       "max_shell_by_size": {str(k): int(v) for k, v in dict(max_shell_by_size).items()},
                       "structure": structure, "symmetrize_decorations": self.symmetrize_decorations}
        L1, L2, L3 = self.supercell
        n_cells = L1 * L2 * L3
        self.n_sites = 2 * n_cells
        c = np.arange(n_cells)
        self.cell_of_site = np.repeat(np.stack([c // (L2 * L3), (c // L3) % L2, c % L3], axis=1), 2, axis=0)
        self.sublattice_of_site = np.tile(np.array([0, 1]), n_cells)
        self.positions = self.cell_of_site @ self.prim + self.basis_pos[self.sublattice_of_site]
        self.active = [len(b) > 1 for b in self.basis_elements]
        if not any(self.active):
            raise ValueError("at least one sublattice needs two species")
        self.background_indices = [int(i) for i in range(self.n_sites) if not self.active[i % 2]]
        self.groups = [b for b in range(2) if self.active[b]]
        self.index_by_trans_symm = [[int(i) for i in range(self.n_sites) if i % 2 == b] for b in self.groups]
        phi = polynomial_basis(len(self.species))
        self.basis = phi
        self.basis_functions = [{sp: float(phi[d, k]) for k, sp in enumerate(self.species)}
                                for d in range(phi.shape[0])]
        self._build(dict(max_shell_by_size))

    def site_of(self, points):
        """Site index of absolute lattice points (a/2 units), wrapped."""
        p = np.asarray(points, dtype=np.int64)
        b = _sublattice(p, self.structure)
        num = (p - self.basis_pos[b]) @ self._adj
        if np.any(num % self._det):
            raise ValueError("point is not a lattice site")
        cell = num // self._det
        L1, L2, L3 = self.supercell
        ci = ((cell[..., 0] % L1) * L2 + (cell[..., 1] % L2)) * L3 + (cell[..., 2] % L3)
        return 2 * ci + b

    def _build(self, max_shell_by_size):
        from itertools import combinations
        info = [dict() for _ in self.groups]
        for gi in range(len(self.groups)):
            info[gi]["c0"] = {"name": "c0", "descriptor": "empty", "size": 0, "ref_indx": self.groups[gi],
                              "symm_group": gi, "max_cluster_dia": 0.0, "indices": [], "order": [],
                              "equiv_sites": []}
            info[gi]["c1"] = {"name": "c1", "descriptor": "point_cluster", "size": 1, "ref_indx": self.groups[gi],
                              "symm_group": gi, "max_cluster_dia": 0.0, "indices": [], "order": [[0]],
                              "equiv_sites": []}
        found = []  # (size, d2, key, group index, absolute clusters, ranks, equiv)
        if max_shell_by_size:
            pts = _crystal_points_within(400, self.structure)
            shells = sorted(set((pts * pts).sum(axis=1).tolist()))
            for size in sorted(max_shell_by_size):
                if size not in (2, 3, 4):
                    raise ValueError("Only cluster sizes 2, 3 and 4 are supported")
                d2_cut = shells[max_shell_by_size[size] - 1]
                for gi, g in enumerate(self.groups):
                    origin = self.basis_pos[g]
                    cand = _crystal_points_within(d2_cut, self.structure) + origin
                    cand = cand[np.array([self.active[int(v)] for v in _sublattice(cand, self.structure)],
                                         dtype=bool)]
                    diff = cand[:, None, :] - cand[None, :, :]
                    ok = (diff * diff).sum(axis=2) <= d2_cut
                    combos = [cmb for cmb in combinations(range(len(cand)), size - 1)
                              if all(ok[a, b] for a, b in combinations(cmb, 2))]
                    if not combos:
                        continue
                    cl = np.concatenate([np.broadcast_to(origin, (len(combos), 1, 3)),
                                         cand[np.array(combos, dtype=np.int64)]], axis=1)
                    rot = np.einsum("rij,cnj->crni", _OPS, cl)
                    codes = _encode(rot)
                    ranks = np.argsort(np.argsort(codes, axis=2, kind="stable"), axis=2, kind="stable")
                    srt_idx = np.argsort(codes, axis=2, kind="stable")
                    srt = np.take_along_axis(codes, srt_idx, axis=2)
                    first_pt = np.take_along_axis(rot, srt_idx[..., :1, None], axis=2)[..., 0, :]
                    sub0 = _sublattice(first_pt, self.structure)  # (n_cl, 48)
                    rel = srt - srt[..., :1] + _CODE_ZERO
                    packed = np.zeros(codes.shape[:2], dtype=np.int64)
                    shift = 0
                    for j in range(1, size):
                        packed += rel[..., j].astype(np.int64) << shift
                        shift += 21
                    # orbit key: smallest (sublattice of the first point, packed)
                    smin = sub0.min(axis=1)
                    big = np.iinfo(np.int64).max
                    masked = np.where(sub0 == smin[:, None], packed, big)
                    pmin = masked.min(axis=1)
                    stab = (sub0 == smin[:, None]) & (packed == pmin[:, None])
                    first_op = np.argmax(stab, axis=1)
                    canon = ranks[np.arange(len(cl)), first_op]
                    d = cl[:, :, None, :] - cl[:, None, :, :]
                    d2max = (d * d).sum(axis=3).reshape(len(cl), -1).max(axis=1)
                    for key in sorted(set(zip(smin.tolist(), pmin.tolist()))):
                        sel = np.nonzero((smin == key[0]) & (pmin == key[1]))[0]
                        rep = sel[0]
                        eq = _equiv_groups(ranks[rep], stab[rep], first_op[rep])
                        found.append((size, int(d2max[rep]), key, gi, cl[sel], canon[sel], eq))
        # global names: orbits of one size by diameter, then key
        keys = sorted({(f[0], f[1], f[2]) for f in found})
        names = {}
        for size in sorted({k[0] for k in keys}):
            same = [k for k in keys if k[0] == size]
            dias = sorted({k[1] for k in same})
            for di, d2 in enumerate(dias):
                for k_i, k in enumerate([k for k in same if k[1] == d2]):
                    names[k] = "c%d_d%04d_%d" % (size, di, k_i)
        self.orbit_names = names
        self.orbit_diameter = {names[k]: math.sqrt(k[1]) / 2.0 for k in keys}
        self.orbit_sizes = {names[k]: k[0] for k in keys}
        self.orbit_equiv = {}
        used = set()
        for size, d2, key, gi, cls, canon, eq in found:
            name = names[(size, d2, key)]
            members, order = [], []
            for m in range(len(cls)):
                by_pos = np.argsort(canon[m])
                r = int(canon[m][0])
                mem = [int(self.site_of(cls[m][k])) for k in by_pos if k != 0]
                members.append(mem)
                order.append([k + 1 if k < r else (0 if k == r else k) for k in range(size)])
                used.update(mem)
            if any(self.groups[gi] in mem or len(set(mem)) != len(mem) for mem in members):
                self.self_interaction = True
            eqs = [list(g) for g in eq] if self.symmetrize_decorations else []
            self.orbit_equiv[name] = eqs
            info[gi][name] = {"name": name, "descriptor": "%d-body d2=%d sub=%d" % (size, d2, key[0]),
                              "size": size, "ref_indx": self.groups[gi], "symm_group": gi,
                              "max_cluster_dia": math.sqrt(d2) / 2.0, "indices": members, "order": order,
                              "equiv_sites": eqs}
        self.cluster_info = info
        # trans(i, col): the site that plays col's role when site i is the
        # reference site of its group: col translated by cell(i)
        cols = np.array(sorted(used), dtype=np.int64)
        self.trans_cols = cols.astype(np.int32)
        shift = self.cell_of_site[:, None, :] @ self.prim  # (n, 1, 3) translation of each site's cell
        tab = self.site_of(self.positions[cols][None, :, :] + shift)
        self.trans_table = np.ascontiguousarray(tab.astype(np.int32))
        self.trans_matrix = _TransMatrixRows(self.trans_cols, self.trans_table)

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


class FCCSettings(CubicSettings):
    """Settings of a periodic FCC supercell (CLEASE ``CEBulk`` with
    ``crystalstructure="fcc"``)."""

    classtype = "FCCSettings"

    def __init__(self, size, species, max_shell_by_size, symmetrize_decorations=True):
        super().__init__(size, species, max_shell_by_size, symmetrize_decorations, "fcc")


class BCCSettings(CubicSettings):
    """Settings of a periodic BCC supercell (``crystalstructure="bcc"``)."""

    classtype = "BCCSettings"

    def __init__(self, size, species, max_shell_by_size, symmetrize_decorations=True):
        super().__init__(size, species, max_shell_by_size, symmetrize_decorations, "bcc")


class SCSettings(CubicSettings):
    """Settings of a periodic simple-cubic supercell (``crystalstructure="sc"``)."""

    classtype = "SCSettings"

    def __init__(self, size, species, max_shell_by_size, symmetrize_decorations=True):
        super().__init__(size, species, max_shell_by_size, symmetrize_decorations, "sc")


def random_crystal_occupation(settings, seed):
    """Species ids of a CrystalSettings lattice, each site drawn uniformly
    from its sublattice's elements (Philox stream of ``seed``)."""
    rng = np.random.Generator(np.random.Philox(int(seed)))
    sp = list(settings.species)
    occ = np.empty(settings.n_sites, dtype=np.int8)
    for i in range(settings.n_sites):
        elems = settings.basis_elements[int(settings.sublattice_of_site[i])]
        occ[i] = sp.index(elems[int(rng.integers(len(elems)))])
    return occ


def seeded_ecis(settings, seed, sigma, scale_by_diameter=True):
    """ECIs drawn from N(0, sigma) in sorted-name order from a Philox stream and
    scaled by 1/diameter^2 for clusters of two or more sites."""
    names = settings.cf_names()
    rng = np.random.Generator(np.random.Philox(seed))
    vals = rng.normal(0.0, sigma, len(names))
    eci = {}
    for name, v in zip(names, vals):
        d = settings.cf_diameter(name)
        if scale_by_diameter and d > 0:
            v = v / (d * d)
        eci[name] = float(v)
    return eci


def random_occupation(n_sites, n_species, seed, fractions=None):
    """Species ids with the given fractions, placed by a seeded permutation."""
    rng = np.random.Generator(np.random.Philox(seed))
    if fractions is None:
        fractions = [1.0 / n_species] * n_species
    counts = [int(round(f * n_sites)) for f in fractions]
    counts[-1] = n_sites - sum(counts[:-1])
    occ = np.concatenate([np.full(c, s, dtype=np.int8) for s, c in enumerate(counts)])
    return occ[rng.permutation(n_sites)]


class Atom:
    __slots__ = ("symbol", "index")

    def __init__(self, symbol, index):
        self.symbol = symbol
        self.index = index


class SimpleAtoms:
    """Minimal atoms container: ``len``, ``atoms[i].symbol`` get/set, iteration."""

    def __init__(self, symbols):
        self._atoms = [Atom(s, i) for i, s in enumerate(symbols)]
        self.calc = None

    def __len__(self):
        return len(self._atoms)

    def __getitem__(self, i):
        return self._atoms[i]

    def __iter__(self):
        return iter(self._atoms)

    def get_chemical_symbols(self):
        return [a.symbol for a in self._atoms]

    def set_calculator(self, calc):
        self.calc = calc

    def get_calculator(self):
        return self.calc

    def copy(self):
        return SimpleAtoms(self.get_chemical_symbols())


def shells_within(diameter, a, lattice="fcc"):
    """Number of neighbour shells of the lattice whose radius is at most
    ``diameter`` for lattice constant ``a`` (shell radii sqrt(d2) * a / 2)."""
    n = 1
    while True:
        sh = lattice_shells(n + 1, lattice)
        if len(sh) <= n or np.sqrt(sh[n]) * a / 2.0 > diameter + 1e-9:
            break
        n += 1
    return n if np.sqrt(lattice_shells(1, lattice)[0]) * a / 2.0 <= diameter + 1e-9 else 0


def fcc_shells_within(diameter, a):
    """Number of FCC neighbour shells whose radius is at most ``diameter``."""
    return shells_within(diameter, a, "fcc")


_BY_NAME = {"fcc": FCCSettings, "bcc": BCCSettings, "sc": SCSettings}


def settings_from_kwargs(kwargs):
    """Settings from the ``setting_kwargs`` of a saved calculator.

    ``classtype`` "FCCSettings", "BCCSettings", "SCSettings",
    "CubicSettings" or "CrystalSettings": this module's own arguments;
    "HCPSettings" or "Spacegroup217Settings": those of cemc_amd.crystal.  ``classtype`` "CEBulk" with
    ``crystalstructure`` "fcc", "bcc" or "sc" (the reference's CLEASE bulk
    settings, ce_calculator.py:577-592): ``size`` is the repetition of the
    primitive cell, the species come from ``basis_elements`` (or
    ``concentration["basis_elements"]``), and each cluster size up to
    ``max_cluster_size`` keeps the neighbour shells within
    ``max_cluster_dia`` (or ``max_cluster_dist``): one diameter for every
    size, or a list over sizes 2, 3, ... (a list of max_cluster_size + 1
    entries is indexed by size).  Other CLEASE settings are not available."""
    kw = dict(kwargs)
    classtype = kw.pop("classtype", "FCCSettings")
    if classtype in ("HCPSettings", "Spacegroup217Settings"):
        from .crystal import HCPSettings, Spacegroup217Settings
        dia = {int(k): float(v) for k, v in kw["max_dia_by_size"].items()}
        if classtype == "HCPSettings":
            return HCPSettings(kw["size"], kw["species"], dia, kw.get("c_over_a", math.sqrt(8.0 / 3.0)),
                               kw.get("symmetrize_decorations", True))
        return Spacegroup217Settings(kw["size"], kw["a"], [tuple(w) for w in kw["wyckoff"]], kw["wyckoff_elements"],
                                     dia, kw.get("symmetrize_decorations", True))
    if classtype == "CrystalSettings":
        shells = {int(k): int(v) for k, v in kw["max_shell_by_size"].items()}
        return CrystalSettings(kw["size"], kw["basis_elements"], shells, kw["structure"],
                               kw.get("symmetrize_decorations", True))
    if classtype in ("FCCSettings", "BCCSettings", "SCSettings", "CubicSettings"):
        shells = {int(k): int(v) for k, v in kw["max_shell_by_size"].items()}
        lattice = kw.get("lattice", {"FCCSettings": "fcc", "BCCSettings": "bcc", "SCSettings": "sc"}.get(classtype))
        return _BY_NAME[str(lattice).lower()](kw["size"], kw["species"], shells, kw.get("symmetrize_decorations", True))
    lattice = str(kw.get("crystalstructure", "")).lower()
    if classtype != "CEBulk" or lattice not in _BY_NAME:
        raise ValueError("Unknown setting classtype: {} (only fcc, bcc and sc bulk settings are available)"
                         .format(kwargs))
    elements = kw.get("basis_elements")
    if elements is None and isinstance(kw.get("concentration"), dict):
        elements = kw["concentration"].get("basis_elements")
    if not elements or len(elements) != 1:
        raise ValueError("Cubic bulk settings need one basis with its elements")
    size = kw["size"]
    if isinstance(size, int):
        size = [size] * 3
    max_size = int(kw.get("max_cluster_size", 4))
    dia = kw.get("max_cluster_dia", kw.get("max_cluster_dist"))
    if dia is None:
        raise ValueError("max_cluster_dia is required")
    a = float(kw["a"])
    if np.isscalar(dia):
        dias = {n: float(dia) for n in range(2, max_size + 1)}
    elif len(dia) == max_size + 1:
        dias = {n: float(dia[n]) for n in range(2, max_size + 1)}
    elif len(dia) == max_size - 1:
        dias = {n: float(dia[n - 2]) for n in range(2, max_size + 1)}
    else:
        raise ValueError("max_cluster_dia must be a number or one value per cluster size")
    shells = {n: shells_within(d, a, lattice) for n, d in dias.items()}
    shells = {n: k for n, k in shells.items() if k > 0}
    return _BY_NAME[lattice](list(size), list(elements[0]), shells)
