"""BCC and simple-cubic lattices on the GPU: the batched kernel (one move at a
time and rounds of eight, closed-form periodic and translation-table paths)
reproduces the oracle's accept sequence, lattice and energy."""
import numpy as np
import pytest

from cemc_amd.lattice import CubicSettings, random_occupation, seeded_ecis
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater

E_TOL = 1e-10  # eV per site

pytestmark = pytest.mark.gpu


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


@pytest.mark.parametrize("lattice", ["bcc", "sc"])
@pytest.mark.parametrize("kind,species", [(0, ["Al", "Mg", "Si"]), (1, ["Al", "Mg"]), (1, ["Al", "Mg", "Si", "Cu"])])
def test_cubic_lattice_matches_oracle(gpu, lattice, kind, species):
    from cemc_amd.batch import PATH_EVALUATE, ChainBatch
    from cemc_amd.model import CEModel
    s = CubicSettings(7, species, {2: 3, 3: 2, 4: 1}, lattice=lattice)
    eci = seeded_ecis(s, 61, 0.01)
    occ = random_occupation(s.n_sites, len(species), 5)
    T, n_moves, seed, uid = 800.0, 4000, 8, 2
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    ref_bits, _, acc = ora.run(kind, seed, uid, n_moves, T)
    ref = _bits(ref_bits, n_moves)
    model = CEModel(s, eci)
    assert model.periodic
    energies = []
    for per_round in (1, 8):
        print(lattice, species, "kind", kind, "per_round", per_round, flush=True)
        batch = ChainBatch([model], [dict(kind=kind, temperature=T, occupation=occ, seed=seed, uid=uid)])
        batch.set_energy_path(PATH_EVALUATE)
        batch.set_moves_per_round(per_round)
        log = batch.run_logged(n_moves)
        mism = np.nonzero(_bits(log[0], n_moves) != ref)[0]
        assert mism.size == 0, "first mismatching move %d" % mism[0]
        assert np.array_equal(batch.occupation(0), ora.get_occupation())
        e = batch.energies()[0]
        assert abs(e - ora.get_energy()) / s.n_sites <= E_TOL
        cf, _ = batch.cf(0)
        assert np.max(np.abs(cf - ora.get_cf_array())) <= 1e-10
        energies.append(e)
        batch.close()
    assert energies[0] == energies[1] and acc > 0


@pytest.mark.parametrize("lattice", ["bcc", "sc"])
def test_cubic_lattice_table_path_matches_oracle(gpu, lattice):
    """A 2x2x3 cell (too small for the closed-form box: clusters wrap onto
    themselves) runs on the translation table."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    s = CubicSettings((2, 2, 3), ["Al", "Mg"], {2: 2, 3: 1}, lattice=lattice)
    eci = seeded_ecis(s, 62, 0.01)
    occ = random_occupation(s.n_sites, 2, 6)
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    ref_bits, _, _ = ora.run(0, 4, 1, 3000, 600.0)
    model = CEModel(s, eci)
    assert not model.periodic
    batch = ChainBatch([model], [dict(kind=0, temperature=600.0, occupation=occ, seed=4, uid=1)])
    log = batch.run_logged(3000)
    assert np.array_equal(_bits(log[0], 3000), _bits(ref_bits, 3000))
    assert np.array_equal(batch.occupation(0), ora.get_occupation())
    assert abs(batch.energies()[0] - ora.get_energy()) / s.n_sites <= E_TOL
