"""Wide rounds (each of 2, 4 or 8 waves of a chain evaluates its own eight
moves of a round) against the reference fixtures, against moves made one at
a time, and against the oracle on large lattices: accept bits, lattices,
energies and sample sums must be bitwise the same as the sequential loop's."""
import numpy as np
import pytest

from golden_io import ChainFixture

pytestmark = pytest.mark.gpu

E_TOL = 1e-10


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


@pytest.mark.parametrize("fname", ["binary_sgc_chain2.npz", "ternary_canonical.npz", "ternary_sgc.npz",
                                   "binary_canonical_plain.npz"])
@pytest.mark.parametrize("waves", [2, 4, 8])
def test_wide_rounds_reproduce_fixtures(gpu, fname, waves):
    """Small cells: most rounds end early, which exercises the cut.  (The
    quaternary fixture's 1746 records per site do not fit one wave, so it
    cannot run wide rounds.)"""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    fx = ChainFixture(fname)
    n = min(fx.n_moves, 20000)
    model = CEModel(fx.settings, fx.eci)
    chain = dict(kind=fx.kind, temperature=fx.temperature, occupation=fx.occ0, seed=fx.seed, uid=fx.chain_uid,
                 chem_pot=fx.meta["mu"] or None)
    batch = ChainBatch([model], [chain])
    batch.set_moves_per_round(8 * waves)
    print("fixture", fname, "wide rounds", waves, "waves", batch.launch_shape(0), flush=True)
    assert batch.launch_shape(0)[0] == waves
    log = batch.run_logged(n)
    assert np.array_equal(_bits(log[0], n), _bits(fx.accept_bits, n))
    if n == fx.n_moves:
        assert np.array_equal(batch.occupation(0), fx.occ_final)
        assert abs(batch.energies()[0] - fx.e_trace[-1]) / fx.n_sites <= E_TOL
    else:
        k = n // fx.trace_every - 1
        assert abs(batch.energies()[0] - fx.e_trace[k]) / fx.n_sites <= E_TOL


def _pair(species, shells, L, kind, temps, n_moves, seed, per_round):
    """The same chains run one move at a time on one wave (the records laid
    out per lane as wide rounds lay them out in every wave) and with
    `per_round`."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    s = FCCSettings(L, species, shells)
    eci = seeded_ecis(s, seed, 0.02)
    model = CEModel(s, eci)
    mu = {"c1_%d" % d: 0.03 * (d + 1) for d in range(len(species) - 1)} if kind == 1 else None
    chains = [dict(kind=kind, temperature=T, occupation=random_occupation(s.n_sites, len(species), seed + c),
                   seed=seed, uid=c, chem_pot=mu) for c, T in enumerate(temps)]
    out = []
    for pr in (1, per_round):
        b = ChainBatch([model] * len(chains), chains)
        b.set_moves_per_round(pr)
        if pr == 1:
            b.set_waves(1)
            assert b.launch_shape(0)[0] == 1
        log = b.run_logged(n_moves)
        out.append((b, log))
    return s, eci, chains, out


@pytest.mark.parametrize("species,shells,L,kind,per_round", [
    (["A", "B"], {2: 4, 3: 2}, 16, 1, 64),
    (["A", "B"], {2: 4, 3: 2}, 16, 0, 32),
    (["A", "B", "C"], {2: 4, 3: 3, 4: 2}, 12, 0, 32),
    (["A", "B", "C"], {2: 3, 3: 2}, 14, 1, 16),
    (["A", "B", "C", "D"], {2: 3, 3: 2}, 12, 0, 64),
])
def test_wide_rounds_equal_sequential_moves(gpu, species, shells, L, kind, per_round):
    temps = [300.0, 900.0, 2500.0]
    n = 6000
    s, eci, chains, ((seq, log_s), (wide, log_w)) = _pair(species, shells, L, kind, temps, n, 17, per_round)
    print("wide vs sequential", len(species), "species L", L, "kind", kind, "per round", per_round,
          "shape", wide.launch_shape(0), flush=True)
    assert wide.launch_shape(0)[0] == per_round // 8
    assert np.array_equal(log_s, log_w)
    assert np.array_equal(seq.energies(), wide.energies())  # bitwise
    for c in range(len(chains)):
        assert np.array_equal(seq.occupation(c), wide.occupation(c))
        ts, tw = seq.thermo(c), wide.thermo(c)
        for key in ("mean_energy", "var_energy"):
            assert ts[key] == tw[key]
        for key in ("singlet", "var_singlet", "cov_singlet_energy"):
            assert np.array_equal(ts[key], tw[key])
    st_s, st_w = seq.stats(), wide.stats()
    for key in st_s:
        assert np.array_equal(st_s[key], st_w[key])
    assert st_w["accepted"].min() > 0


def test_wide_rounds_match_oracle_on_a_large_lattice(gpu):
    """Binary FCC 24^3 (13824 sites, the largest serving cell), canonical,
    wide rounds of 32 against the oracle's chain on the same stream."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    s = FCCSettings(24, ["A", "B"], {2: 4, 3: 2})
    eci = seeded_ecis(s, 2024, 0.02)
    occ = random_occupation(s.n_sites, 2, 5)
    batch = ChainBatch([CEModel(s, eci)], [dict(kind=0, temperature=700.0, occupation=occ, seed=3, uid=0)])
    batch.set_moves_per_round(32)
    n = 20000
    log = batch.run_logged(n)
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    bits, _, acc = ora.run(0, 3, 0, n, 700.0)
    assert np.array_equal(_bits(log[0], n), _bits(bits, n))
    assert np.array_equal(batch.occupation(0), ora.get_occupation())
    assert abs(batch.energies()[0] - ora.get_energy()) / s.n_sites <= E_TOL
    assert batch.stats()["accepted"][0] == acc
