"""GPU parity against the oracle at the benchmarked lattice sizes, where the
size-dependent parts of the kernel (padded-box LDS layout, the canonical
tracker of up to 32768 entries, energy tables in HBM, several waves per chain,
wide rounds) run as they do in the bench:

* ternary FCC 16^3 (4096 sites), the largest cell of the uneven-sizes batch,
  with its cluster set, on the default shape, on rounds of eight over four
  waves and on wide rounds;
* quaternary FCC 32^3 (32768 sites), the tempering workload's model: two
  replicas, 2000 swaps each in two exchange cycles, against the oracle's
  parallel tempering;
* binary and ternary FCC 24^3 (13824 sites), the largest serving cells,
  canonical and SGC, in one batch;
* jobs of a serving run (sizes 6^3-24^3) against oracle chains of the same
  jobs.

Accept bits are compared bit for bit; energies within 1e-10 eV/site.
"""
import numpy as np
import pytest

from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater

pytestmark = pytest.mark.gpu

E_TOL = 1e-10


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


def _tile(small, ls, L):
    """Occupation of an L^3 cell repeating an ls^3 cell (site = (x L + y) L + z)."""
    x, y, z = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    return np.asarray(small)[((x % ls) * ls + y % ls) * ls + z % ls].reshape(-1).astype(np.int8)


def _oracle(settings, names, occ, eci, mu=None, cf0=None):
    e = dict(eci)
    for k, v in (mu or {}).items():
        e[k] -= v
    cf = cf0 if cf0 is not None else brute_force_cf(settings, occ, names)
    return OracleUpdater(OracleModel(settings, names), occ, cf, e)


def test_crowded_ternary_launch_reads_tables_from_hbm(gpu):
    """1100 ternary FCC 8^3 chains with the uneven-sizes cluster set: more
    than can be resident at once with their energy tables in LDS, so the
    launch reads the tables from HBM (one 16-byte global load per record)
    and runs more chains per compute unit; chains equal oracle chains."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    st = FCCSettings(8, ["A", "B", "C"], {2: 4, 3: 3, 4: 2})
    eci = seeded_ecis(st, 2026, 0.01)
    names = st.cf_names()
    n_chains = 1100
    temps = np.geomspace(300.0, 1500.0, n_chains)
    occs = [random_occupation(st.n_sites, 3, 7000 + c) for c in range(n_chains)]
    chains = [dict(kind=CANONICAL, temperature=float(temps[c]), occupation=occs[c], seed=501, uid=c)
              for c in range(n_chains)]
    batch = ChainBatch([CEModel(st, eci)] * n_chains, chains)
    shape = batch.launch_shape(0)
    print("crowded ternary launch shape", shape, flush=True)
    assert not shape[2], "tables expected in HBM"
    n = 2000
    log = batch.run_logged(n)
    for c in (0, 549, 1099):
        ora = _oracle(st, names, occs[c], eci)
        bits, _, acc = ora.run(0, 501, c, n, temps[c])
        assert np.array_equal(_bits(log[c], n), _bits(bits, n)), c
        assert np.array_equal(batch.occupation(c), ora.get_occupation())
        assert abs(batch.energies()[c] - ora.get_energy()) / st.n_sites <= E_TOL
        assert batch.stats()["accepted"][c] == acc


@pytest.mark.parametrize("shape", ["auto", "4x8", "wide"])
def test_ternary_16_cube_matches_oracle(gpu, shape):
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    st = FCCSettings(16, ["A", "B", "C"], {2: 4, 3: 3, 4: 2})
    eci = seeded_ecis(st, 2025, 0.01)
    names = st.cf_names()
    temps = [450.0, 1300.0]
    occs = [random_occupation(st.n_sites, 3, 5000 + c, fractions=f)
            for c, f in enumerate(([0.5, 0.3, 0.2], [0.2, 0.2, 0.6]))]
    chains = [dict(kind=CANONICAL, temperature=T, occupation=o, seed=500, uid=c)
              for c, (T, o) in enumerate(zip(temps, occs))]
    batch = ChainBatch([CEModel(st, eci)] * 2, chains)
    if shape == "4x8":
        batch.set_moves_per_round(8)
        batch.set_waves(4)
    elif shape == "wide":
        batch.set_moves_per_round(32)
    n = 3000
    print("ternary 16^3", shape, "launch shape", batch.launch_shape(0), flush=True)
    log = batch.run_logged(n)
    for c in range(2):
        ora = _oracle(st, names, occs[c], eci)
        bits, _, acc = ora.run(0, 500, c, n, temps[c])
        assert np.array_equal(_bits(log[c], n), _bits(bits, n))
        assert np.array_equal(batch.occupation(c), ora.get_occupation())
        assert abs(batch.energies()[c] - ora.get_energy()) / st.n_sites <= E_TOL
        assert batch.stats()["accepted"][c] == acc


def test_quaternary_32_cube_tempering_matches_oracle(gpu):
    """The tempering workload's model at its size (tables in HBM, eight waves
    per chain, rounds of eight, a 36^3 padded box and a 32768-entry tracker
    in LDS): two replicas, two cycles of 1000 swaps and an exchange each.
    Starting lattices repeat an 8^3 cell, whose brute-force correlation
    functions are those of the 32^3 cell (every site sees the same
    neighbourhood), so the GPU's from-scratch CFs are checked at 32^3 too."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    from cemc_amd.tempering import ReplicaExchange
    from oracle.pt import run_pt
    shells = {2: 3, 3: 3, 4: 3}
    st = FCCSettings(32, ["A", "B", "C", "D"], shells)
    small = FCCSettings(8, ["A", "B", "C", "D"], shells)
    eci = seeded_ecis(st, 2026, 0.02)
    names = st.cf_names()
    assert names == small.cf_names()
    occ8 = [random_occupation(small.n_sites, 4, 7000 + r) for r in range(2)]
    occs = [_tile(o, 8, 32) for o in occ8]
    cf8 = [brute_force_cf(small, o, names) for o in occ8]
    temps = [400.0, 1600.0]
    chains = [dict(kind=CANONICAL, temperature=T, occupation=occs[r], seed=9, uid=r) for r, T in enumerate(temps)]
    batch = ChainBatch([CEModel(st, eci)] * 2, chains)
    w, kr, glds = batch.launch_shape(0)
    print("quaternary 32^3 launch shape", (w, kr, glds), flush=True)
    assert w == 8 and not glds
    for r in range(2):
        cf, _ = batch.cf(r)
        assert np.max(np.abs(cf - np.array([cf8[r][n] for n in names]))) <= 1e-12
    pt = ReplicaExchange(batch, seed=9, uid=1 << 40)
    pt.run(2, 1000, use_graph=True)
    stp = pt.state()
    ups = [_oracle(st, names, occs[r], eci, cf0=cf8[r]) for r in range(2)]
    cos, n_acc, samples = run_pt(ups, 0, [9, 9], [0, 1], temps, 2, 1000, 9, 1 << 40)
    assert np.array_equal(stp["chain_of_slot"], cos)
    assert np.array_equal(stp["n_accept"], n_acc)
    for r in range(2):
        assert np.array_equal(batch.occupation(r), ups[r].get_occupation())
        assert abs(batch.energies()[r] - ups[r].get_energy()) / st.n_sites <= E_TOL
        assert abs(stp["mean_energy"][r] - samples[r].mean()) <= 1e-9 * st.n_sites
    assert batch.stats()["accepted"].min() > 0


def test_serving_size_cells_match_oracle(gpu):
    """Binary and ternary 24^3 cells, canonical and SGC, in one batch
    (launch shapes chosen automatically: four chains leave room for wide
    rounds)."""
    from cemc_amd.batch import CANONICAL, SGC, ChainBatch
    from cemc_amd.model import CEModel
    bin_s = FCCSettings(24, ["A", "B"], {2: 4, 3: 2})
    ter_s = FCCSettings(24, ["A", "B", "C"], {2: 4, 3: 3, 4: 2})
    systems = [(bin_s, seeded_ecis(bin_s, 2024, 0.02)), (ter_s, seeded_ecis(ter_s, 2027, 0.01))]
    models = {id(s): CEModel(s, e) for s, e in systems}
    jobs = []
    for k, (s, e) in enumerate(systems):
        n_sp = len(s.species)
        for kind in (CANONICAL, SGC):
            mu = {"c1_%d" % d: 0.02 * (d + 1) for d in range(n_sp - 1)} if kind == SGC else None
            jobs.append((s, e, kind, mu, 600.0 + 400.0 * len(jobs), random_occupation(s.n_sites, n_sp, 60 + len(jobs))))
    chains = [dict(kind=kind, temperature=T, occupation=o, seed=77, uid=j, chem_pot=mu)
              for j, (s, e, kind, mu, T, o) in enumerate(jobs)]
    batch = ChainBatch([models[id(j[0])] for j in jobs], chains)
    n = 5000
    print("24^3 serving cells, shapes", [batch.launch_shape(c) for c in range(len(jobs))], flush=True)
    log = batch.run_logged(n)
    for j, (s, e, kind, mu, T, o) in enumerate(jobs):
        ora = _oracle(s, s.cf_names(), o, e, mu)
        bits, _, acc = ora.run(kind, 77, j, n, T)
        assert np.array_equal(_bits(log[j], n), _bits(bits, n)), j
        assert np.array_equal(batch.occupation(j), ora.get_occupation())
        assert abs(batch.energies()[j] - ora.get_energy()) / s.n_sites <= E_TOL


def test_served_jobs_match_oracle_chains(gpu):
    """A serving run (continuous batching, jobs admitted as they arrive):
    completed jobs of every size class equal oracle chains of the same jobs
    (move kind, temperature, chemical potentials, stream), not only the same
    jobs run alone on the GPU."""
    from cemc_amd.batch import CANONICAL, SGC
    from cemc_amd.model import CEModel
    from cemc_amd.serving import Job, JobServer
    rng = np.random.Generator(np.random.Philox(12))
    cache = {}
    jobs = []
    for i, (L, ternary, kind) in enumerate([(6, False, SGC), (24, False, CANONICAL), (10, True, CANONICAL),
                                             (20, True, SGC), (14, False, SGC), (8, True, CANONICAL),
                                             (24, True, CANONICAL), (12, False, CANONICAL)]):
        species = ["A", "B", "C"] if ternary else ["A", "B"]
        shells = {2: 4, 3: 3, 4: 2} if ternary else {2: 4, 3: 2}
        key = (L, ternary)
        if key not in cache:
            st = FCCSettings(L, species, shells)
            eci = seeded_ecis(st, 2027 if ternary else 2024, 0.01 if ternary else 0.02)
            cache[key] = (st, eci, CEModel(st, eci))
        st, eci, model = cache[key]
        mu = {"c1_%d" % d: float(rng.uniform(-0.1, 0.1)) for d in range(len(species) - 1)} if kind == SGC else None
        sweeps = 1 if (ternary and L >= 20) else 3
        jobs.append(Job(job_id=i, arrival=0.01 * i, model=model, kind=kind, temperature=float(rng.uniform(300, 1500)),
                        sweeps=sweeps, occupation=random_occupation(st.n_sites, len(species), 300 + i), seed=77,
                        uid=i, chem_pot=mu))
    server = JobServer(max_chains=64, hbm_bytes=64 << 20, quantum=2.0e6)
    res = server.serve(jobs, realtime=False)
    assert res["jobs"] == len(jobs) and not server.rejected
    print("served", res["jobs"], "jobs in", res["launches"], "launches", flush=True)
    for job in server.done:
        L = int(round(job.model.n_sites ** (1.0 / 3.0)))
        st, eci, _ = cache[(L, job.model.n_species == 3)]
        ora = _oracle(st, st.cf_names(), job.occupation, eci, job.chem_pot)
        _, _, acc = ora.run(job.kind, job.seed, job.uid, job.n_moves, job.temperature)
        assert np.array_equal(job.final_occupation, ora.get_occupation()), job.job_id
        assert abs(job.energy - ora.get_energy()) / job.model.n_sites <= E_TOL
        assert job.accepted == acc
