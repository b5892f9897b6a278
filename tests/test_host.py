"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point its header declares, invalid model descriptions are
rejected before any device work, the package refuses to run without its HIP
library, and the lattice/cluster settings satisfy the geometric and algebraic
facts the cluster expansion relies on."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from cemc_amd import _capi
from cemc_amd.lattice import (FCCSettings, enumerate_orbits, equivalent_decorations, fcc_points_within,
                              polynomial_basis, random_occupation, seeded_ecis)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cemc_gpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cemc_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_match_binding_and_library_exports():
    declared = _declared()
    assert len(declared) >= 25
    assert sorted(n for n, _, _ in _capi.SIGNATURES) == declared
    lib = _capi.load()
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.library_path()], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (cemc_[a-z0-9_]+)", out))
    assert set(declared) <= exported
    # the library exports C symbols only (no C++ names leak through the ABI)
    assert not [s for s in re.findall(r"\bT (\S+)", out) if s.startswith("_Z") and "cemc" in s]


def test_library_has_gfx950_code_object():
    # the library's offload bundle names its only target
    data = open(_capi.library_path(), "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "_LIB_PATH", os.path.join(ROOT, "no_such_dir", "libcemcgpu.so"))
    with pytest.raises(ImportError):
        _capi.load()


def _no_gpu():
    return _capi.device_count() == 0


@pytest.mark.skipif(not os.path.exists(_capi.library_path()), reason="library not built")
def test_no_gpu_is_reported_not_emulated():
    if not _no_gpu():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="GPU"):
        _capi.require_gpu()


def _minimal_desc(**over):
    """A one-group binary model with a single nearest-neighbour pair entry on
    a 2-site ring, as raw arrays; fields can be overridden to make it invalid."""
    keep = []

    def i32(v):
        a, p = _capi.i32(np.array(v))
        keep.append(a)
        return p

    def f64(v):
        a, p = _capi.f64(np.array(v, dtype=float))
        keep.append(a)
        return p

    f = dict(n_sites=2, n_species=2, n_groups=1, n_cf=3, n_entries=1, n_trans_cols=1,
             basis=[1.0, -1.0], site_group=[0, 0], cf_kind=[0, 1, 2], cf_dec=[0, 0, 0], cf_entry=[-1, -1, 0],
             cf_eq_off=[0, 0, 0, 1], eq_decos=[0, 0, 0, 0], entry_group=[0], entry_size=[2], entry_M=[1],
             entry_norm=[1], entry_mem_off=[0, 1], members=[0], entry_ord_off=[0, 2], order=[0, 1],
             trans=[1, 0], eci=[0.0, 0.1, 0.2], background=[0, 0])
    f.update(over)
    d = _capi.ModelDesc()
    for k, v in f.items():
        if isinstance(v, list):
            setattr(d, k, f64(v) if k in ("basis", "eci") else i32(v))
        else:
            setattr(d, k, v)
    return d, keep


@pytest.mark.parametrize("over,msg", [
    (dict(n_species=9), "n_species"),
    (dict(n_groups=0), "n_groups"),
    (dict(site_group=[0, 3]), "group out of range"),
    (dict(site_group=[0, -1]), "not been assigned"),
    (dict(trans=[1, 2]), "translation table"),
    (dict(entry_size=[5]), "cluster sizes"),
    (dict(order=[0, 0]), "permutation"),
    (dict(cf_kind=[0, 1, 7]), "kind"),
    (dict(members=[4]), "member column"),
    (dict(background=[1, 1]), "no movable site"),
])
def test_invalid_model_descriptions_are_rejected(over, msg):
    lib = _capi.load()
    d, keep = _minimal_desc(**over)
    h = ctypes.c_void_p()
    rc = lib.cemc_model_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _capi.CEMC_EINVAL
    assert msg in lib.cemc_last_error().decode()
    assert not h.value


def test_null_arguments_are_rejected():
    lib = _capi.load()
    assert lib.cemc_model_create(None, None) == _capi.CEMC_EINVAL
    assert lib.cemc_batch_run(None, None, 0) == _capi.CEMC_EINVAL
    assert lib.cemc_updater_update_cf(None, 0, 0, 1) == _capi.CEMC_EINVAL


# ----------------------------------------------------------------- lattice


def test_fcc_neighbour_shells():
    p = fcc_points_within(10)
    d2 = (p * p).sum(axis=1)
    assert [int((d2 == v).sum()) for v in (2, 4, 6, 8, 10)] == [12, 6, 24, 12, 24]


def test_orbit_multiplicities():
    orbits = {o.name: o for o in enumerate_orbits({2: 2, 3: 1, 4: 1})}
    # clusters containing a given FCC site: 12 nn pairs, 6 nnn pairs,
    # 24 nn triangles, 8 nn tetrahedra
    assert orbits["c2_d0000_0"].multiplicity == 12
    assert orbits["c2_d0001_0"].multiplicity == 6
    assert orbits["c3_d0000_0"].multiplicity == 24
    assert orbits["c4_d0000_0"].multiplicity == 8
    for o in orbits.values():
        # every sub-cluster holds the reference site at the origin
        for cl in o.offsets:
            assert any(not np.any(v) for v in cl)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_basis_is_orthonormal(n):
    phi = polynomial_basis(n)
    assert phi.shape == (n - 1, n)
    np.testing.assert_allclose(phi @ phi.T / n, np.eye(n - 1), atol=1e-13)
    np.testing.assert_allclose(phi.sum(axis=1), 0.0, atol=1e-13)


def test_translation_table_is_a_lattice_translation():
    s = FCCSettings((4, 5, 6), ["A", "B"], {2: 3, 3: 1})
    L = np.array(s.dims)
    for j, c in enumerate(s.trans_cols):
        shift = s.trans_col_cells[j]
        want = s.site_index((s.cells + shift) % L)
        assert np.array_equal(s.trans_table[:, j], want)
        # column c is the image of the reference site 0
        assert s.trans_table[0, j] == c


def test_cluster_info_consistency():
    s = FCCSettings(6, ["Al", "Mg", "Si"], {2: 3, 3: 2, 4: 1})
    info = s.cluster_info[0]
    for name, cl in info.items():
        n = cl["size"]
        if n < 2:
            continue
        assert len(cl["indices"]) == len(cl["order"])
        for sub, order in zip(cl["indices"], cl["order"]):
            assert len(sub) == n - 1 and sorted(order) == list(range(n))
            assert cl["ref_indx"] not in sub
    names = s.cf_names()
    assert names[0] == "c0" and names[1:3] == ["c1_0", "c1_1"]
    assert names == sorted(names)


def test_equivalent_decorations_are_permutations():
    """Decorations equivalent under permutations of equivalent cluster
    positions (the reference calls ase.clease.tools.equivalent_deco, an
    unpinned ``ase`` dependency absent from the reference tree,
    cpp/src/cluster.cpp:77-123).  The list holds every permutation, so each
    distinct decoration appears equally often (orbit-stabilizer): the average
    the updater takes over it (cpp/src/ce_updater.cpp:391-400) is the same
    with or without repeats."""
    import itertools
    from collections import Counter
    decos = equivalent_decorations([0, 0, 1], [[0, 1, 2]])
    assert len(decos) == 6
    counts = Counter(map(tuple, decos))
    assert set(counts) == {(0, 0, 1), (0, 1, 0), (1, 0, 0)} and len(set(counts.values())) == 1
    decos = equivalent_decorations([0, 1, 2, 1], [[1, 2, 3]])
    counts = Counter(map(tuple, decos))
    want = {(0,) + p for p in itertools.permutations((1, 2, 1))}
    assert set(counts) == want and len(set(counts.values())) == 1
    assert equivalent_decorations([0, 1], []) == [[0, 1]]


def test_seeded_inputs_are_deterministic():
    s = FCCSettings(4, ["A", "B"], {2: 2})
    assert seeded_ecis(s, 3, 0.01) == seeded_ecis(s, 3, 0.01)
    a = random_occupation(s.n_sites, 2, 9)
    assert np.array_equal(a, random_occupation(s.n_sites, 2, 9))
    assert a.dtype == np.int8 and set(np.unique(a)) <= {0, 1}
