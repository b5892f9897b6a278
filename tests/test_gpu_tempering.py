"""Parallel tempering on the GPU against the oracle's restatement on the same
fed stream: the slot permutation, the exchange acceptance counts, every
lattice and the per-temperature energy averages must agree, with and without
the captured HIP graph."""
import numpy as np
import pytest

from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater
from oracle.pt import run_pt

pytestmark = pytest.mark.gpu


def _setup(species, shells, size, R, seed=5):
    s = FCCSettings(size, species, shells)
    eci = seeded_ecis(s, seed, 0.02)
    occs = [random_occupation(s.n_sites, len(species), 100 + r) for r in range(R)]
    return s, eci, occs


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("species,shells,size,R", [
    (["A", "B"], {2: 4, 3: 2}, 6, 6),
    (["Al", "Mg", "Si"], {2: 3, 3: 2, 4: 1}, 5, 5),
])
def test_tempering_matches_oracle(gpu, use_graph, species, shells, size, R):
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    from cemc_amd.tempering import ReplicaExchange
    s, eci, occs = _setup(species, shells, size, R)
    temps = np.geomspace(250.0, 2500.0, R).tolist()
    model = CEModel(s, eci)
    chains = [dict(kind=CANONICAL, temperature=T, occupation=occs[r], seed=11, uid=r) for r, T in enumerate(temps)]
    batch = ChainBatch([model] * R, chains)
    pt = ReplicaExchange(batch, seed=21, uid=99)
    n_cycles, moves = 40, s.n_sites // 2
    print("replicas", R, "cycles", n_cycles, "graph", use_graph, flush=True)
    pt.run(n_cycles - 7, moves, use_graph=use_graph)
    pt.run(7, moves, use_graph=use_graph)
    st = pt.state()

    names = s.cf_names()
    om = OracleModel(s, names)
    ups = [OracleUpdater(om, o, brute_force_cf(s, o, names), eci) for o in occs]
    cos, n_acc, samples = run_pt(ups, 0, [11] * R, list(range(R)), temps, n_cycles, moves, 21, 99)
    assert np.array_equal(st["chain_of_slot"], cos)
    assert np.array_equal(st["n_accept"], n_acc)
    assert st["n_try"].sum() == n_cycles * (R // 2)
    for c in range(R):
        assert np.array_equal(batch.occupation(c), ups[c].get_occupation())
        assert abs(batch.energies()[c] - ups[c].get_energy()) <= 1e-10 * s.n_sites
    for t in range(R):
        assert st["n_samples"][t] == len(samples[t])
        assert abs(st["mean_energy"][t] - samples[t].mean()) <= 1e-9
        assert abs(st["var_energy"][t] - samples[t].var()) <= 1e-7 * max(1.0, samples[t].var())
    # temperatures follow their slots
    assert np.allclose(st["temperatures"], temps, rtol=1e-15)
    assert n_acc.sum() > 0


def test_parallel_tempering_facade(gpu):
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.mcmc import Montecarlo
    from cemc_amd.tempering import ParallelTempering
    s, eci, occs = _setup(["A", "B"], {2: 2}, 6, 1)
    atoms = SimpleAtoms([s.species[v] for v in occs[0]])
    CE(atoms, s, eci=dict(eci))
    mc = Montecarlo(atoms, 800.0, seed=3)
    pt = ParallelTempering(mc, Tmax=2000.0, Tmin=200.0, temp_scheme_file="", n_replicas=8)
    pt.run(mc_args={"steps": s.n_sites}, num_exchange_cycles=20)
    res = pt.get_thermodynamic()
    assert len(res) == 8
    assert all(r["n_samples"] == 20 * s.n_sites for r in res)
    # lower temperature, lower mean energy (on average over the run)
    assert res[-1]["energy"] <= res[0]["energy"] + 1e-9
    occ = pt.replica_occupation(0)
    assert occ.shape == (s.n_sites,)


def _oracle_scheme(s, eci, occ, kind, seed, uid0, Tmax, Tmin, target=0.2):
    """The reference's scheme search (parallel_tempering.py:47-136) driven by
    oracle chains on the same streams: replica r is a new chain with uid
    uid0 + r started from the previous replica's lattice; every trial is a
    runMC that resets the averages first (montecarlo.py:796-800), so each
    mean covers that trial's moves only."""
    from oracle.pt import accept_probability
    names = s.cf_names()
    model = OracleModel(s, names)
    n = 10 * s.n_sites

    def new(o):
        return {"upd": OracleUpdater(model, o, brute_force_cf(s, o, names), eci), "t": 0}

    def run(ch, T, r):
        _, tr, _ = ch["upd"].run(kind, seed, uid0 + r, n, T, move_begin=ch["t"], trace_every=1)
        ch["t"] += n
        ch["E"] = float(np.mean(tr))
        return ch["E"]

    cur = new(occ)
    run(cur, Tmax, 0)
    T_cur, temps, accs, r = Tmax, [Tmax], [0.0], 1
    while T_cur > Tmin:
        cur_E = cur["E"]
        nw = new(cur["upd"].get_occupation())
        trial, acc, found = T_cur / 2.0, 1.0, False
        while acc > target and trial > Tmin:
            acc = accept_probability(cur_E, run(nw, trial, r), T_cur, trial)
            found = acc <= target
            trial /= 2.0
        if not found:
            break
        hi, lo = T_cur, trial
        while True:
            t = 0.5 * (hi + lo)
            acc = accept_probability(cur_E, run(nw, t, r), T_cur, t)
            if acc > target:
                hi = t
            else:
                lo = t
            if abs(acc - target) < 0.01 or hi - lo < 1e-4:
                break
        temps.append(t)
        accs.append(acc)
        cur, T_cur, r = nw, t, r + 1
    return temps, accs


def test_temperature_scheme_search_matches_oracle(gpu):
    """ParallelTempering without a scheme runs the reference's search on the
    GPU; the same search on oracle chains with the same streams picks the
    same temperatures; consecutive replicas meet the acceptance target."""
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.mcmc import Montecarlo
    from cemc_amd.tempering import ParallelTempering, find_temperature_scheme
    s = FCCSettings(4, ["Al", "Mg"], {2: 2, 3: 1})
    eci = seeded_ecis(s, 8, 0.02)
    occ = random_occupation(s.n_sites, 2, 4)
    atoms = SimpleAtoms([s.species[v] for v in occ])
    CE(atoms, s, eci=dict(eci))
    mc = Montecarlo(atoms, 1000.0, seed=17, uid=40)
    temps, accs = find_temperature_scheme(mc, 2000.0, 150.0)
    ref_t, ref_a = _oracle_scheme(s, eci, occ, 0, 17, 40, 2000.0, 150.0)
    print("scheme", [round(t, 3) for t in temps], flush=True)
    assert len(temps) >= 3 and len(temps) == len(ref_t)
    assert np.allclose(temps, ref_t, rtol=0, atol=1e-9) and np.allclose(accs, ref_a, rtol=0, atol=1e-9)
    assert all(a > b for a, b in zip(temps, temps[1:]))
    pt = ParallelTempering(mc, Tmax=2000.0, Tmin=150.0, temp_scheme_file="")
    assert np.allclose(pt.temperature_scheme, temps, rtol=0, atol=1e-9)
    pt.run(mc_args={"steps": s.n_sites}, num_exchange_cycles=4)
