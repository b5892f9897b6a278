"""The parallel-tempering oracle (oracle/pt.py) pinned on the CPU by what the
mathematics fixes, on a 2x2x3 FCC binary cell whose 924 fixed-composition
states are enumerated (19 energy levels):

* lattices frozen (no moves between exchanges): the slot permutation is a
  Markov chain whose stationary law is exactly
  pi(sigma) ~ exp(-sum_t E_sigma(t) / kT_t) over the 3! permutations;
* with sweeps between exchanges: every slot samples the Boltzmann law of its
  own temperature, and each pair's exchange acceptance equals the exact
  expectation of min(1, exp((1/kT_a - 1/kT_b)(E_a - E_b))) under the
  product of the two slots' laws;
* the pairs tried follow the reference's up/down lists
  (parallel_tempering.py:153-160): with an even replica count the pairs
  (1,2), (3,4), ... are never tried, with an odd count alternate pairs are
  tried equally often.

Planted mistakes in the oracle (a flipped sign of d(1/kT), pairs that ignore
the direction) fail these checks, which shows they bite.
"""
import itertools

import numpy as np
import pytest

from cemc_amd.lattice import KB_EV, FCCSettings
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater
from oracle.pt import accept_probability, exchange_pairs, run_pt

TEMPS = [500.0, 900.0, 1600.0]


@pytest.fixture(scope="module")
def cell():
    s = FCCSettings((2, 2, 3), ["A", "B"], {2: 3})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    pairs = [n for n in names if n.startswith("c2")]
    for n, v in zip(pairs, (0.012, -0.02, 0.007)):
        eci[n] = v
    N = s.n_sites
    states = [np.array(st, dtype=np.int8) for st in itertools.product([0, 1], repeat=N) if sum(st) == N // 2]
    energies = np.array([N * sum(eci[n] * v for n, v in brute_force_cf(s, st, names).items()) for st in states])
    levels, level_of = np.unique(np.round(energies, 9), return_inverse=True)
    return dict(s=s, names=names, eci=eci, states=states, energies=energies, levels=levels, level_of=level_of,
                model=OracleModel(s, names))


def _updater(cell, occ):
    return OracleUpdater(cell["model"], occ, brute_force_cf(cell["s"], occ, cell["names"]), cell["eci"])


def _boltzmann_levels(cell, T):
    e = cell["energies"]
    w = np.exp(-(e - e.min()) / (KB_EV * T))
    p = np.zeros(len(cell["levels"]))
    np.add.at(p, cell["level_of"], w)
    return p / p.sum()


def _permutation_law(energies, temps):
    perms = list(itertools.permutations(range(len(temps))))
    logw = np.array([-sum(energies[sig[t]] / (KB_EV * temps[t]) for t in range(len(temps))) for sig in perms])
    w = np.exp(logw - logw.max())
    return perms, w / w.sum()


def _frozen_run(cell, n_cycles, **planted):
    # three lattices about 0.1 eV apart (every permutation keeps > 2 % weight)
    idx = [int(np.argmin(np.abs(cell["energies"] - e))) for e in (-0.126, -0.023, 0.065)]
    ups = [_updater(cell, cell["states"][i]) for i in idx]
    energies = [u.get_energy() for u in ups]
    rec = {}
    run_pt(ups, 0, [3] * 3, [0, 1, 2], TEMPS, n_cycles, 0, 77, 5, record=rec, **planted)
    perms, law = _permutation_law(energies, TEMPS)
    counts = {p: 0 for p in perms}
    for h in rec["history"]:
        counts[h] += 1
    freq = np.array([counts[p] for p in perms]) / len(rec["history"])
    return freq, law, rec


def test_frozen_lattices_sample_the_exact_permutation_law(cell):
    freq, law, rec = _frozen_run(cell, 60000)
    assert law.min() > 0.01  # every permutation matters
    assert np.abs(freq - law).max() < 0.02  # planted mistakes below miss by > 0.25
    assert rec["n_try"].sum() == 60000  # one pair per cycle with three replicas


def test_frozen_law_check_catches_a_flipped_sign(cell):
    flipped = lambda E1, E2, T1, T2: accept_probability(E2, E1, T1, T2)  # noqa: E731
    freq, law, _ = _frozen_run(cell, 6000, accept=flipped)
    assert np.abs(freq - law).max() > 0.05


def test_frozen_law_check_catches_pairs_that_ignore_the_direction(cell):
    only_up = lambda R, up: exchange_pairs(R, True)  # noqa: E731
    freq, law, _ = _frozen_run(cell, 6000, pairs=only_up)
    assert np.abs(freq - law).max() > 0.05


def _sweep_run(cell, n_cycles, moves, **planted):
    occs = [cell["states"][i] for i in (0, 300, 600)]
    ups = [_updater(cell, o) for o in occs]
    rec = {}
    cos, n_acc, samples = run_pt(ups, 0, [11] * 3, [0, 1, 2], TEMPS, n_cycles, moves, 21, 9, record=rec, **planted)
    lv = {round(float(v), 9): k for k, v in enumerate(cell["levels"])}
    shares = []
    for t in range(3):
        h = np.zeros(len(cell["levels"]))
        for e in samples[t]:
            h[lv[round(float(e), 9)]] += 1
        shares.append(h / h.sum())
    return shares, n_acc, rec


def _exact_exchange_rate(cell, Ta, Tb):
    pa, pb = _boltzmann_levels(cell, Ta), _boltzmann_levels(cell, Tb)
    L = cell["levels"]
    r = 0.0
    for i, ea in enumerate(L):
        for j, eb in enumerate(L):
            r += pa[i] * pb[j] * min(1.0, accept_probability(ea, eb, Ta, Tb))
    return r


def test_slots_sample_boltzmann_and_exchange_rates_are_exact(cell):
    shares, n_acc, rec = _sweep_run(cell, 12000, 12)
    for t, T in enumerate(TEMPS):
        want = _boltzmann_levels(cell, T)
        assert 0.5 * np.abs(shares[t] - want).sum() < 0.03, (T, shares[t], want)
    for k in range(2):
        rate = n_acc[k] / rec["n_try"][k]
        exact = _exact_exchange_rate(cell, TEMPS[k], TEMPS[k + 1])
        assert 0.1 < exact < 0.9
        assert abs(rate - exact) < 0.03, (k, rate, exact)


def test_slot_law_check_catches_a_flipped_sign(cell):
    flipped = lambda E1, E2, T1, T2: accept_probability(E2, E1, T1, T2)  # noqa: E731
    shares, _, _ = _sweep_run(cell, 3000, 12, accept=flipped)
    tv = [0.5 * np.abs(shares[t] - _boltzmann_levels(cell, T)).sum() for t, T in enumerate(TEMPS)]
    assert max(tv) > 0.1


@pytest.mark.parametrize("R", [4, 5])
def test_pairs_follow_the_reference_lists(cell, R):
    occs = [cell["states"][i] for i in range(0, 100 * R, 100)]
    ups = [_updater(cell, o) for o in occs]
    temps = list(np.geomspace(500.0, 1600.0, R))
    rec = {}
    run_pt(ups, 0, [2] * R, list(range(R)), temps, 400, 0, 31, 4, record=rec)
    t = rec["n_try"]
    if R == 4:
        # up: (0,1), (2,3); down: (3,2), (1,0) -- the pair (1,2) is never tried
        assert t[1] == 0 and t[0] == t[2] == 400
    else:
        # up: (0,1), (2,3); down: (4,3), (2,1); the direction is a fair coin
        assert t[0] + t[1] == 400 and 140 < t[0] < 260
        assert t[0] == t[2] and t[1] == t[3]
    assert exchange_pairs(5, True) == [(0, 1), (2, 3)] and exchange_pairs(5, False) == [(4, 3), (2, 1)]
