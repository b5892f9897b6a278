"""Settings of crystals whose sites are not all related by translations
(cemc_amd.crystal): hexagonal close packing and a body-centred cubic
crystal of space group I-43m (217) with four Wyckoff positions, the
reference's own test structures (tests/test_CE_updater.py:21-59 runs
"hcp"; :206-228 and tests/helper_functions.py:8-27 the space-group-217
binary).  CPU checks against geometry and against brute-force correlation
functions:

* the space groups found from the structure have the expected orders
  (P6_3/mmc: 24 operations modulo translations; I-43m: 24 x 2 with the
  centring);
* HCP pair orbits are its neighbour shells, with the textbook coordination
  numbers 12 (6 in-plane + 6 out-of-plane), 6, 2, 18 and every member at
  the orbit's distance;
* space group 217: Wyckoff multiplicities 2, 8, 24, 24 per conventional
  cell, 29 translation groups of two sites each (the centring), pair members
  at the orbit's distance;
* the oracle's incremental correlation functions equal brute force after
  random flips and swaps (binary and ternary HCP, binary 217);
* HCP binary energy changes equal the Ising local field built from Cartesian
  neighbour vectors alone;
* the settings rebuild from their saved arguments.
"""
import itertools

import numpy as np
import pytest

from cemc_amd.crystal import HCPSettings, Spacegroup217Settings, hcp_cell, space_group_ops
from cemc_amd.lattice import random_occupation, seeded_ecis, settings_from_kwargs
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater

WYCKOFF_217 = [(0, 0, 0), (0.324, 0.324, 0.324), (0.3582, 0.3582, 0.0393), (0.0954, 0.0954, 0.2725)]


def _min_image(st, i, j):
    """Minimum-image Cartesian vector from site i to site j."""
    box = np.array(st.supercell)[:, None] * st.cell
    f = ((st.cell_of_site[j] + st.frac[st.unit_of_site[j]]) - (st.cell_of_site[i] + st.frac[st.unit_of_site[i]])) \
        / np.array(st.supercell)
    f -= np.round(f)
    best = None
    for k in itertools.product((-1, 0, 1), repeat=3):
        v = (f + np.array(k)) @ box
        if best is None or v @ v < best @ best:
            best = v
    return best


def test_space_groups_found_from_the_structures():
    cell, frac = hcp_cell()
    assert len(space_group_ops(cell, frac, [0, 0])) == 24
    st = Spacegroup217Settings(1, 10.553, WYCKOFF_217, [["Al", "Mg"]] * 4, {2: 4.1})
    assert len(st.ops) == 48
    assert st.n_sites == 58 and len(st.index_by_trans_symm) == 29
    assert all(len(g) == 2 for g in st.index_by_trans_symm)
    counts = np.bincount([int(np.argmin([np.linalg.norm(np.mod(f - np.array(w) + 0.5, 1.0) - 0.5) for w in WYCKOFF_217]))
                          for f in st.frac[:1]], minlength=1)
    assert counts.sum() == 1
    # Wyckoff multiplicities from the expansion (2a, 8c, 24g, 24g)
    src = [len({tuple(np.round(f, 6)) for f in st.frac if _orbit_of(f) == r}) for r in range(4)]
    assert src == [2, 8, 24, 24]


def _orbit_of(f):
    from cemc_amd.crystal import tetrahedral_point_ops
    for r, w in enumerate(WYCKOFF_217):
        for R in tetrahedral_point_ops():
            for t in (np.zeros(3), np.full(3, 0.5)):
                d = np.mod(R @ np.array(w) + t - f + 0.5, 1.0) - 0.5
                if np.all(np.abs(d) < 1e-6):
                    return r
    return -1


def test_hcp_pair_orbits_are_its_neighbour_shells():
    st = HCPSettings(5, ["Al", "Mg"], {2: 1.75})
    c = st.cell[2, 2]
    shells = {}
    for g, info in enumerate(st.cluster_info):
        ref = info["c0"]["ref_indx"]
        for name, cl in info.items():
            if cl["size"] != 2:
                continue
            for (j,) in cl["indices"]:
                v = _min_image(st, ref, j)
                assert abs(np.linalg.norm(v) - cl["max_cluster_dia"]) < 1e-9
                key = round(float(np.linalg.norm(v)), 6)
                shells.setdefault((g, key), []).append(abs(v[2]) < 1e-9)
    for g in range(2):
        assert len(shells[(g, 1.0)]) == 12 and sum(shells[(g, 1.0)]) == 6
        assert len(shells[(g, round(np.sqrt(2.0), 6))]) == 6
        assert len(shells[(g, round(c, 6))]) == 2
        assert len(shells[(g, round(np.sqrt(3.0), 6))]) == 18
    # the in-plane and out-of-plane nearest neighbours are different orbits
    nn = sorted(name for name, cl in st.cluster_info[0].items()
                if cl["size"] == 2 and abs(cl["max_cluster_dia"] - 1.0) < 1e-9)
    assert len(nn) == 2


def test_spacegroup_217_pair_members_at_their_distance():
    st = Spacegroup217Settings(1, 10.553, WYCKOFF_217, [["Al", "Mg"]] * 4, {2: 4.1, 3: 4.1})
    assert not st.self_interaction
    for info in st.cluster_info:
        ref = info["c0"]["ref_indx"]
        for cl in info.values():
            if cl["size"] == 2:
                for (j,) in cl["indices"]:
                    assert abs(np.linalg.norm(_min_image(st, ref, j)) - cl["max_cluster_dia"]) < 1e-6


@pytest.mark.parametrize("make,species", [
    (lambda sp: HCPSettings(4, sp, {2: 1.75, 3: 1.45}), ["Al", "Mg"]),
    (lambda sp: HCPSettings((4, 4, 3), sp, {2: 1.45, 3: 1.05, 4: 1.05}), ["Al", "Mg", "Si"]),
    (lambda sp: Spacegroup217Settings(1, 10.553, WYCKOFF_217, [sp] * 4, {2: 4.1, 3: 4.1}), ["Al", "Mg"]),
])
def test_incremental_cf_equals_bruteforce(make, species):
    st = make(species)
    names = st.cf_names()
    n_s = len(species)
    eci = seeded_ecis(st, 4, 0.01)
    occ = random_occupation(st.n_sites, n_s, 8)
    upd = OracleUpdater(OracleModel(st, names), occ, brute_force_cf(st, occ, names), eci)
    rng = np.random.Generator(np.random.Philox(12))
    cur = occ.copy()
    for step in range(30):
        i, j = (int(v) for v in rng.integers(0, st.n_sites, 2))
        if step % 2 and cur[i] != cur[j]:
            chs = [(i, int(cur[i]), int(cur[j])), (j, int(cur[j]), int(cur[i]))]
        else:
            chs = [(i, int(cur[i]), int((cur[i] + 1 + step % (n_s - 1)) % n_s))]
        e = upd.calculate(chs)
        upd.clear_history()
        cur = upd.get_occupation()
        ref = brute_force_cf(st, cur, names)
        ref_arr = np.array([ref[n] for n in names])
        assert np.max(np.abs(upd.get_cf_array() - ref_arr)) < 1e-12
        assert abs(e - st.n_sites * float(np.dot([eci[n] for n in names], ref_arr))) < 1e-10


def test_hcp_binary_energy_change_equals_ising_local_field():
    """Pairs to the 4th HCP shell; each pair orbit identified from its
    Cartesian vector alone (length, and in-plane or not)."""
    st = HCPSettings(5, ["A", "B"], {2: 1.75})
    names = st.cf_names()
    eci = seeded_ecis(st, 19, 0.02)
    occ = random_occupation(st.n_sites, 2, 23)
    upd = OracleUpdater(OracleModel(st, names), occ, brute_force_cf(st, occ, names), eci)
    # orbit of each geometric class, from the reference site of group 0
    ref0 = st.cluster_info[0]["c0"]["ref_indx"]
    cls_of = {}
    for name, cl in st.cluster_info[0].items():
        if cl["size"] == 2:
            v = _min_image(st, ref0, cl["indices"][0][0])
            cls_of[(round(float(np.linalg.norm(v)), 6), abs(v[2]) < 1e-9)] = name
    e0 = upd.get_energy()
    rng = np.random.Generator(np.random.Philox(2))
    for i in (int(v) for v in rng.integers(0, st.n_sites, 20)):
        now = upd.get_occupation()
        spin = np.where(now == 0, 1.0, -1.0)
        field = {}
        count = {}
        for j in range(st.n_sites):
            if j == i:
                continue
            v = _min_image(st, i, j)
            key = (round(float(np.linalg.norm(v)), 6), abs(v[2]) < 1e-9)
            if key in cls_of:
                field[key] = field.get(key, 0.0) + spin[j]
                count[key] = count.get(key, 0) + 1
        # E = N J cf with cf the mean over the N M / 2 pairs of an orbit:
        # flipping s_i changes E by J (-2 s_i) (2 / M) sum_j s_j
        closed = eci["c1_0"] * (-2.0 * spin[i])
        for key, name in cls_of.items():
            closed += eci[name + "_00"] * (-2.0 * spin[i]) * 2.0 / count[key] * field[key]
        e1 = upd.calculate([(i, int(now[i]), 1 - int(now[i]))])
        assert abs((e1 - e0) - closed) < 1e-12
        upd.undo_changes()


def test_settings_rebuild_from_saved_arguments():
    for st in (HCPSettings(3, ["Al", "Mg"], {2: 1.45}),
               Spacegroup217Settings(1, 10.553, WYCKOFF_217, [["Al", "Mg"]] * 4, {2: 4.1})):
        again = settings_from_kwargs(dict(st.kwargs, classtype=type(st).__name__))
        assert again.cf_names() == st.cf_names()
        assert np.array_equal(again.trans_table, st.trans_table)
