"""get_atoms_with_ce_calc (cemc/ce_calculator.py:48-102; the reference's
test is tests/test_CE_updater.py:117-157): a large cell's calculator starts
from the correlation functions of a small cell's template.

CPU: the fact the helper relies on -- the template cells of one crystal at
two sizes have the same brute-force correlation functions -- and the
argument checks that run before any GPU call.  GPU: the large cell's
calculator follows brute force through single-site changes, as the
reference's test checks."""
import json

import numpy as np
import pytest

from cemc_amd.calculator import get_atoms_with_ce_calc, get_max_size_eci, template_symbols
from cemc_amd.crystal import HCPSettings
from cemc_amd.lattice import CrystalSettings, FCCSettings, seeded_ecis
from oracle.brute_force import brute_force_cf


def _template_cf(st):
    sp = list(st.species)
    occ = np.array([sp.index(s) for s in template_symbols(st)], dtype=np.int8)
    return brute_force_cf(st, occ, st.cf_names())


@pytest.mark.parametrize("make", [
    lambda n: FCCSettings(n, ["Al", "Mg"], {2: 3, 3: 2}),
    lambda n: CrystalSettings(n, [["Na", "K"], ["Cl", "Br"]], {2: 2, 3: 1}),
    lambda n: HCPSettings(n, ["Al", "Mg", "Si"], {2: 1.45, 3: 1.05}),
])
def test_template_cells_share_correlation_functions(make):
    small, large = make(4), make(5)
    a, b = _template_cf(small), _template_cf(large)
    assert a.keys() == b.keys()
    for k in a:
        assert abs(a[k] - b[k]) < 1e-12, k


def test_template_symbols_follow_sublattices():
    st = CrystalSettings(2, [["Na", "K"], ["Cl"]], {2: 2})
    sym = template_symbols(st)
    assert all(s == ("K" if sub == 0 else "Cl") for s, sub in zip(sym, st.sublattice_of_site))


def test_eci_larger_than_tracked_clusters_is_rejected():
    st = FCCSettings(4, ["Al", "Mg"], {2: 3})
    assert get_max_size_eci({"c0": 0.0, "c1_0": 1.0, "c3_d0000_0_000": 0.1}) == 3
    with pytest.raises(ValueError, match="larger than"):
        get_atoms_with_ce_calc(st, st.kwargs, {"c3_d0000_0_000": 0.1}, size=[5, 5, 5])


@pytest.mark.gpu
@pytest.mark.parametrize("classtype", ["FCCSettings", "CEBulk"])
def test_supercell_calculator_follows_bruteforce(gpu, tmp_path, classtype):
    from cemc_amd.calculator import get_atoms_with_ce_calc_JSON
    from cemc_amd.lattice import settings_from_kwargs
    if classtype == "CEBulk":
        kwargs = {"classtype": "CEBulk", "crystalstructure": "fcc", "a": 4.05, "size": [3, 3, 3],
                  "basis_elements": [["Al", "Mg"]], "max_cluster_size": 3, "max_cluster_dia": 4.1}
    else:
        kwargs = dict(FCCSettings(4, ["Al", "Mg"], {2: 3, 3: 2}).kwargs, classtype="FCCSettings")
    small = settings_from_kwargs(kwargs)
    eci = seeded_ecis(small, 3, 0.02)
    path = tmp_path / "settings.json"
    path.write_text(json.dumps(kwargs))
    for atoms in (get_atoms_with_ce_calc(small, {k: v for k, v in kwargs.items() if k != "classtype"}, eci,
                                         size=[6, 6, 6]),
                  get_atoms_with_ce_calc_JSON(str(path), eci, size=[6, 6, 6])):
        calc = atoms.get_calculator()
        large = calc.BC
        assert large.n_sites == 216 and len(atoms) == 216
        names = list(calc.get_cf().keys())
        sp = list(large.species)
        for i in range(10):
            calc.calculate(atoms, ["energy"], [(i, "Al", "Mg")])
            occ = np.array([sp.index(a.symbol) for a in atoms], dtype=np.int8)
            ref = brute_force_cf(large, occ, [n for n in names if n != "c0"])
            got = calc.get_cf()
            for k, v in ref.items():
                assert abs(got[k] - v) < 1e-10, (i, k)
        calc.updater.close()
