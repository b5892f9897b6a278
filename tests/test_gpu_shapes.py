"""GPU parity of every launch shape: one to eight waves per chain, energy
tables in LDS or in HBM, periodic closed-form and translation-table lattices,
and batches that mix models, move kinds and sizes.  Each chain must reproduce
its reference fixture bit for bit whatever shape and company it runs in."""
import numpy as np
import pytest

from golden_io import ChainFixture

pytestmark = pytest.mark.gpu

E_TOL = 1e-10


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


def _chain(fx, **extra):
    d = dict(kind=fx.kind, temperature=fx.temperature, occupation=fx.occ0, seed=fx.seed, uid=fx.chain_uid,
             chem_pot=fx.meta["mu"] or None)
    d.update(extra)
    return d


def _table_settings(settings):
    class Table:
        pass

    t = Table()
    for key, val in vars(settings).items():
        if key not in ("dims", "cells"):
            setattr(t, key, val)
    return t


def _check(fx, batch, c, log):
    ref = _bits(fx.accept_bits, fx.n_moves)
    got = _bits(log[c], fx.n_moves)
    mism = np.nonzero(got != ref)[0]
    assert mism.size == 0, "first mismatching move %d" % mism[0]
    assert np.array_equal(batch.occupation(c), fx.occ_final)
    e = batch.energies()[c]
    assert abs(e - fx.e_trace[-1]) / fx.n_sites <= E_TOL


@pytest.mark.parametrize("fname", ["ternary_canonical.npz", "ternary_sgc.npz", "quaternary_canonical.npz"])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
@pytest.mark.parametrize("per_round", [1, 8])
def test_waves_per_chain(gpu, fname, waves, per_round):
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture(fname)
    print("fixture", fname, "waves", waves, "moves per round", per_round, flush=True)
    model = CEModel(fx.settings, fx.eci)
    batch = ChainBatch([model], [_chain(fx)])
    batch.set_moves_per_round(per_round)
    batch.set_waves(waves)
    w, kr, glds = batch.launch_shape(0)
    print("shape", w, kr, glds, flush=True)
    log = batch.run_logged(fx.n_moves)
    _check(fx, batch, 0, log)


@pytest.mark.parametrize("waves", [1, 4])
def test_translation_table_lattice_with_waves(gpu, waves):
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture("ternary_canonical.npz")
    model = CEModel(_table_settings(fx.settings), fx.eci)
    assert not model.periodic
    batch = ChainBatch([model], [_chain(fx)])
    batch.set_waves(waves)
    log = batch.run_logged(fx.n_moves)
    _check(fx, batch, 0, log)


def test_mixed_batch_matches_each_fixture(gpu):
    """Binary SGC, ternary canonical/SGC and quaternary canonical chains in one
    batch (several launch classes), each run as a whole and in two halves."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    names = ["binary_sgc_chain0.npz", "binary_sgc_chain2.npz", "ternary_canonical.npz", "ternary_sgc.npz",
             "quaternary_canonical.npz"]
    fxs = [ChainFixture(n) for n in names]
    models = [CEModel(fx.settings, fx.eci) for fx in fxs]
    chains = [_chain(fx) for fx in fxs]
    batch = ChainBatch(models, chains)
    # initial energies (energy tables) equal N J.cf from the CF kernel
    for c, fx in enumerate(fxs):
        _, e_cf = batch.cf(c)
        assert abs(batch.energies()[c] - e_cf) <= 1e-12 * fx.n_sites
    n_moves = np.array([fx.n_moves for fx in fxs], dtype=np.int64)
    log = batch.run_logged(n_moves)
    for c, fx in enumerate(fxs):
        _check(fx, batch, c, log)
    # the same chains advanced in two calls of uneven length
    batch2 = ChainBatch(models, chains)
    first = n_moves // 3
    log_a = batch2.run_logged(first)
    log_b = batch2.run_logged(n_moves - first)
    for c, fx in enumerate(fxs):
        a = _bits(log_a[c], int(first[c]))
        b = _bits(log_b[c], int(n_moves[c] - first[c]))
        assert np.array_equal(np.concatenate([a, b]), _bits(fx.accept_bits, fx.n_moves))
        assert np.array_equal(batch2.occupation(c), fx.occ_final)


def test_resumed_chain_continues_its_stream(gpu):
    """A chain re-created from its occupation with first_move = t continues
    exactly where the first batch stopped (canonical tracker rebuilt in
    ascending order, as the reference's tracker is after set_symbols)."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture("binary_sgc_chain1.npz")
    model = CEModel(fx.settings, fx.eci)
    half = fx.n_moves // 2
    b1 = ChainBatch([model], [_chain(fx)])
    log1 = b1.run_logged(half)
    occ_mid = b1.occupation(0)
    b2 = ChainBatch([model], [_chain(fx, occupation=occ_mid, first_move=half)])
    log2 = b2.run_logged(fx.n_moves - half)
    got = np.concatenate([_bits(log1[0], half), _bits(log2[0], fx.n_moves - half)])
    assert np.array_equal(got, _bits(fx.accept_bits, fx.n_moves))
    assert np.array_equal(b2.occupation(0), fx.occ_final)


@pytest.mark.parametrize("fname", ["binary_canonical_plain.npz", "binary_sgc_chain0.npz", "binary_sgc_chain3.npz"])
def test_sequential_and_speculative_rounds_agree(gpu, fname):
    """Binary chains with one move per round and with eight: identical accept
    sequences (both equal to the reference) and energies."""
    from cemc_amd.batch import PATH_EVALUATE, ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture(fname)
    model = CEModel(fx.settings, fx.eci)
    logs, es = [], []
    for per_round in (1, 8):
        batch = ChainBatch([model], [_chain(fx)])
        batch.set_energy_path(PATH_EVALUATE)
        batch.set_moves_per_round(per_round)
        logs.append(batch.run_logged(fx.n_moves))
        _check(fx, batch, 0, logs[-1])
        es.append(batch.energies()[0])
    assert np.array_equal(logs[0], logs[1])
    assert es[0] == es[1]


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("waves", [1, 2, 8])
def test_many_records_per_lane_match_oracle(gpu, kind, waves):
    """Ternary FCC 8^3 with triplets to the 3rd and quadruplets to the 2nd
    shell (494 sub-clusters per site: 8 records per lane on one wave, 4 on
    two, 2 on eight): one move per round and eight give the oracle's accept
    sequence, lattice and energy, and the same energy bitwise."""
    from cemc_amd.batch import PATH_EVALUATE, ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    s = FCCSettings(8, ["A", "B", "C"], {2: 4, 3: 3, 4: 2})
    eci = seeded_ecis(s, 2027, 0.01)
    occ = random_occupation(s.n_sites, 3, 9)
    T, n_moves, seed, uid = 700.0, 4000, 5, 3
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    ref_bits, _, _ = ora.run(kind, seed, uid, n_moves, T)
    ref = _bits(ref_bits, n_moves)
    model = CEModel(s, eci)
    energies = []
    for per_round in (1, 8):
        print("kind", kind, "waves", waves, "per_round", per_round, flush=True)
        batch = ChainBatch([model], [dict(kind=kind, temperature=T, occupation=occ, seed=seed, uid=uid)])
        batch.set_energy_path(PATH_EVALUATE)
        batch.set_moves_per_round(per_round)
        batch.set_waves(waves)
        assert batch.launch_shape(0)[:2] == (waves, {1: 8, 2: 4, 8: 2}[waves])
        log = batch.run_logged(n_moves)
        got = _bits(log[0], n_moves)
        mism = np.nonzero(got != ref)[0]
        assert mism.size == 0, "first mismatching move %d" % mism[0]
        assert np.array_equal(batch.occupation(0), ora.get_occupation())
        e = batch.energies()[0]
        assert abs(e - ora.get_energy()) / s.n_sites <= E_TOL
        energies.append(e)
    assert energies[0] == energies[1]
    assert ref.sum() > 0


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("species,shells,size,T", [
    (["A", "B"], {2: 4, 3: 2}, 10, 600.0),
    (["A", "B", "C"], {2: 4, 3: 3, 4: 2}, 8, 700.0),
    (["A", "B", "C", "D"], {2: 3, 3: 2}, 7, 900.0),
])
def test_fields_match_oracle_and_evaluation(gpu, kind, species, shells, size, T):
    """Per-site fields updated on accepted moves: the oracle's accept
    sequence and lattice, energies within rounding of the oracle's and of the
    evaluating path; swaps of interacting sites included (small lattices)."""
    from cemc_amd.batch import PATH_EVALUATE, PATH_FIELDS, ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    s = FCCSettings(size, species, shells)
    eci = seeded_ecis(s, 77, 0.02)
    occ = random_occupation(s.n_sites, len(species), 12)
    n_moves, seed, uid = 6000, 9, 4
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    ref_bits, _, _ = ora.run(kind, seed, uid, n_moves, T)
    ref = _bits(ref_bits, n_moves)
    model = CEModel(s, eci)
    energies = {}
    for path in (PATH_FIELDS, PATH_EVALUATE):
        print("species", len(species), "kind", kind, "path", path, flush=True)
        batch = ChainBatch([model], [dict(kind=kind, temperature=T, occupation=occ, seed=seed, uid=uid)])
        batch.set_energy_path(path)
        assert batch.energy_path(0) == path
        log = batch.run_logged(n_moves // 2)
        log2 = batch.run_logged(n_moves - n_moves // 2)  # fields rebuilt at the next launch
        got = np.concatenate([_bits(log[0], n_moves // 2), _bits(log2[0], n_moves - n_moves // 2)])
        mism = np.nonzero(got != ref)[0]
        assert mism.size == 0, "first mismatching move %d" % mism[0]
        assert np.array_equal(batch.occupation(0), ora.get_occupation())
        energies[path] = batch.energies()[0]
        assert abs(energies[path] - ora.get_energy()) / s.n_sites <= E_TOL
        st = batch.stats()
        assert st["accepted"][0] == int(ref.sum())
    assert abs(energies[PATH_FIELDS] - energies[PATH_EVALUATE]) / s.n_sites <= 1e-12
    assert ref.sum() > 0


def test_energy_path_choice(gpu):
    """Evaluation by default; fields when asked for on closed-form periodic
    lattices, not with more waves per chain forced nor on translation-table
    lattices."""
    from cemc_amd.batch import PATH_EVALUATE, PATH_FIELDS, ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture("binary_sgc_chain0.npz")
    model = CEModel(fx.settings, fx.eci)
    batch = ChainBatch([model], [_chain(fx)])
    assert batch.energy_path(0) == PATH_EVALUATE
    batch.set_energy_path(PATH_FIELDS)
    assert batch.energy_path(0) == PATH_FIELDS
    batch.set_waves(2)
    assert batch.energy_path(0) == PATH_EVALUATE
    batch.set_waves(0)
    assert batch.energy_path(0) == PATH_FIELDS
    table = CEModel(_table_settings(fx.settings), fx.eci)
    tb = ChainBatch([table], [_chain(fx)])
    tb.set_energy_path(PATH_FIELDS)
    assert tb.energy_path(0) == PATH_EVALUATE


@pytest.mark.parametrize("species", [["A", "B"], ["A", "B", "C"]])
@pytest.mark.parametrize("shells", [{2: 4}, {2: 4, 3: 2}, {2: 4, 3: 3}, {2: 2, 4: 2}, {2: 4, 3: 3, 4: 2}])
def test_fields_equal_evaluation_across_cluster_sets(gpu, species, shells):
    """The field path against the evaluating path over 30 launches of 64
    moves (fields rebuilt per launch, updated within it): identical accept
    sequences and energies within rounding, for flips and swaps, with 2 to 8
    records per lane (ternary swaps with triplets to the 3rd shell included)."""
    from cemc_amd.batch import PATH_EVALUATE, PATH_FIELDS, ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    st = FCCSettings(8, species, shells)
    model = CEModel(st, seeded_ecis(st, 5, 0.02))
    for kind in (0, 1):
        chain = dict(kind=kind, temperature=1200.0, occupation=random_occupation(st.n_sites, len(species), 3),
                     seed=2, uid=1)
        out = {}
        for path in (PATH_EVALUATE, PATH_FIELDS):
            b = ChainBatch([model], [chain])
            b.set_energy_path(path)
            assert b.energy_path(0) == path
            logs = [_bits(b.run_logged(64)[0], 64) for _ in range(30)]
            out[path] = (np.concatenate(logs), b.energies()[0], b.occupation(0))
        assert np.array_equal(out[PATH_EVALUATE][0], out[PATH_FIELDS][0]), (kind, shells)
        assert np.array_equal(out[PATH_EVALUATE][2], out[PATH_FIELDS][2])
        assert abs(out[PATH_EVALUATE][1] - out[PATH_FIELDS][1]) / st.n_sites <= 1e-12
