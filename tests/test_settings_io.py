"""Settings from saved calculator arguments: this package's cubic and
two-sublattice settings and the reference's CLEASE bulk arguments
(ce_calculator.py:573-592)."""
import itertools

import numpy as np
import pytest

from cemc_amd.lattice import FCCSettings, fcc_shells_within, settings_from_kwargs


def test_fcc_shell_count_from_diameter():
    """FCC shells lie at a*sqrt(k/2) for k = 1, 2, 3, 4 (a/sqrt 2, a, a*sqrt 1.5,
    a*sqrt 2): with a = 4.05 a diameter of 4.5 keeps two shells, 4.05 keeps two
    (boundary included), 5.8 keeps four, 2.0 none."""
    a = 4.05
    assert fcc_shells_within(4.5, a) == 2
    assert fcc_shells_within(4.05, a) == 2
    assert fcc_shells_within(5.8, a) == 4
    assert fcc_shells_within(a * np.sqrt(1.5), a) == 3
    assert fcc_shells_within(2.0, a) == 0


def test_cebulk_fcc_arguments():
    kw = {"classtype": "CEBulk", "crystalstructure": "fcc", "a": 4.05, "size": [3, 4, 5],
          "concentration": {"basis_elements": [["Mg", "Al"]]}, "max_cluster_size": 3, "max_cluster_dia": 4.5}
    s = settings_from_kwargs(kw)
    assert s.dims == (3, 4, 5) and s.species == ["Al", "Mg"]
    assert s.kwargs["max_shell_by_size"] == {"2": 2, "3": 2}
    s2 = settings_from_kwargs(dict(kw, max_cluster_dia=[5.8, 4.5]))
    assert s2.kwargs["max_shell_by_size"] == {"2": 4, "3": 2}
    s3 = settings_from_kwargs(dict(kw, max_cluster_dia=[0, 0, 5.8, 4.5]))
    assert s3.kwargs["max_shell_by_size"] == {"2": 4, "3": 2}
    assert s2.cf_names() == s3.cf_names()
    with pytest.raises(ValueError):
        settings_from_kwargs(dict(kw, crystalstructure="hcp"))
    with pytest.raises(ValueError):
        settings_from_kwargs({"classtype": "CECrystal"})


def test_cebulk_bcc_and_sc_arguments():
    """BCC shells at a*sqrt(3)/2, a, a*sqrt(2), a*sqrt(11)/2 and SC shells at
    a, a*sqrt(2), a*sqrt(3): the diameters map to those shell counts, and
    the settings round-trip through their own arguments."""
    from cemc_amd.lattice import BCCSettings, SCSettings, shells_within
    a = 3.0
    assert shells_within(a * np.sqrt(3) / 2, a, "bcc") == 1
    assert shells_within(a * 1.01, a, "bcc") == 2
    assert shells_within(a * np.sqrt(2), a, "bcc") == 3
    assert shells_within(a * np.sqrt(11) / 2, a, "bcc") == 4
    assert shells_within(a * 0.99, a, "sc") == 0
    assert shells_within(a * np.sqrt(2), a, "sc") == 2
    assert shells_within(a * np.sqrt(3), a, "sc") == 3
    kw = {"classtype": "CEBulk", "crystalstructure": "bcc", "a": a, "size": [4, 4, 4],
          "basis_elements": [["Fe", "Cr"]], "max_cluster_size": 3, "max_cluster_dia": [a * 1.01, a * 0.9]}
    b = settings_from_kwargs(kw)
    assert isinstance(b, BCCSettings) and b.n_sites == 64
    assert sorted(n.rsplit("_", 1)[0] for n in b.cf_names() if n.startswith("c2")) == ["c2_d0000_0", "c2_d0001_0"]
    s = settings_from_kwargs(dict(kw, crystalstructure="sc", size=5, basis_elements=[["Cu", "Au", "Ag"]]))
    assert isinstance(s, SCSettings) and s.n_sites == 125
    for st in (b, s):
        again = settings_from_kwargs(dict(st.kwargs, classtype=type(st).__name__))
        assert type(again) is type(st) and again.cf_names() == st.cf_names()
        assert np.array_equal(again.trans_table, st.trans_table)


def test_fcc_settings_round_trip():
    s = FCCSettings(4, ["C", "A", "B"], {2: 3, 3: 2}, symmetrize_decorations=False)
    kw = dict(s.kwargs, classtype="FCCSettings")
    r = settings_from_kwargs(kw)
    assert r.dims == s.dims and r.species == s.species
    assert r.cf_names() == s.cf_names()
    assert r.symmetrize_decorations is False


def test_two_sublattice_crystals():
    """Rocksalt and CsCl: one translation group per active sublattice, the
    nearest-neighbour (unlike-sublattice) pair orbit shared by both groups
    with 6 (rocksalt) or 8 (CsCl) members, like-sublattice orbits named
    per sublattice; a single-element sublattice is background; the members
    of every group's pair orbits lie at the orbit's distance; the settings
    round-trip through their own arguments."""
    from cemc_amd.lattice import CrystalSettings
    rs = CrystalSettings(3, [["Na", "K"], ["Cl", "Br"]], {2: 3}, "rocksalt")
    assert rs.n_sites == 54 and len(rs.index_by_trans_symm) == 2 and rs.background_indices == []
    nn = [{k: len(v["indices"]) for k, v in g.items() if k.startswith("c2")} for g in rs.cluster_info]
    assert nn[0]["c2_d0000_0"] == 6 and nn[1]["c2_d0000_0"] == 6
    assert "c2_d0001_0" in nn[0] and "c2_d0001_0" not in nn[1] and "c2_d0001_1" in nn[1]
    cs = CrystalSettings(4, [["Cu", "Zn"], ["Cu", "Zn"]], {2: 1}, "cscl")
    assert all(len(g["c2_d0000_0"]["indices"]) == 8 for g in cs.cluster_info)
    bg = CrystalSettings(3, [["Na", "K"], ["Cl"]], {2: 2}, "rocksalt")
    assert len(bg.index_by_trans_symm) == 1 and bg.background_indices == list(range(1, 54, 2))
    assert bg.species == ["Cl", "K", "Na"]
    for st in (rs, cs, bg):
        pos = st.positions.astype(float)
        box = np.array(st.supercell)[:, None] * st.prim.astype(float)
        inv = np.linalg.inv(box)
        for g, info in enumerate(st.cluster_info):
            ref = info["c0"]["ref_indx"]
            for name, cl in info.items():
                if cl["size"] != 2:
                    continue
                for (j,) in cl["indices"]:
                    f = (pos[j] - pos[ref]) @ inv
                    f -= np.round(f)
                    # minimum image over the neighbouring supercells (skewed box)
                    d2 = min(float((((f + np.array(k)) @ box) ** 2).sum())
                             for k in itertools.product((-1, 0, 1), repeat=3))
                    assert abs(np.sqrt(d2) / 2.0 - cl["max_cluster_dia"]) < 1e-12
                    assert j not in st.background_indices
        again = settings_from_kwargs(dict(st.kwargs, classtype="CrystalSettings"))
        assert again.cf_names() == st.cf_names() and np.array_equal(again.trans_table, st.trans_table)
