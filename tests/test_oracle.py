"""CPU oracle pinned against the reference and against mathematics.

* every golden fixture written by the reference CEUpdater (tests/golden) is
  reproduced bit for bit: accept/reject bits, energy trace, correlation
  functions, final occupation;
* the reference's recorded updater call log (updater_api_ternary.json) is
  replayed exactly, including the background-site error;
* Philox4x64-10 matches numpy's Philox and the Random123 known answer;
* the incremental correlation functions equal a from-scratch evaluation on
  random flip/swap sequences, with and without equivalent sites;
* binary energy changes equal the Ising local-field closed form built
  directly from lattice geometry;
* exhaustive enumeration on a tiny cell: every flip of every state;
* undo restores bit-identical state; swaps conserve composition.
"""
import itertools
import math

import numpy as np
import pytest

from cemc_amd.lattice import CubicSettings, FCCSettings, FCC_PRIM, KB_EV, random_occupation, seeded_ecis
from golden_io import CHAIN_FIXTURES, CRYSTAL_FIXTURES, ChainFixture, load_api_fixture
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater, philox4x64

MASK64 = (1 << 64) - 1


def _numpy_philox_block(ctr, key):
    c = (sum(v << (64 * i) for i, v in enumerate(ctr)) - 1) % (1 << 256)
    start = np.array([(c >> (64 * i)) & MASK64 for i in range(4)], dtype=np.uint64)
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=start)
    return [int(v) for v in bg.random_raw(4)]


def test_philox_matches_numpy_and_random123():
    # Random123 known answer for philox4x64-10 with zero counter and key
    assert philox4x64((0, 0, 0, 0), (0, 0)) == [0x16554d9eca36314c, 0xdb20fe9d672d0fdc,
                                                 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    rng = np.random.Generator(np.random.Philox(5))
    for _ in range(50):
        ctr = [int(v) for v in rng.integers(0, 2 ** 63, 4)]
        key = [int(v) for v in rng.integers(0, 2 ** 63, 2)]
        assert philox4x64(ctr, key) == _numpy_philox_block(ctr, key)


@pytest.mark.parametrize("fname", CHAIN_FIXTURES + CRYSTAL_FIXTURES)
def test_oracle_reproduces_reference_fixture(fname):
    fx = ChainFixture(fname)
    model = OracleModel(fx.settings, fx.names)
    upd = OracleUpdater(model, fx.occ0, fx.cf0, fx.eci_run)
    bits, cks, trace = [], [], []
    per = fx.cf_every
    for seg in range(fx.n_moves // per):
        b, tr, _ = upd.run(fx.kind, fx.seed, fx.chain_uid, per, fx.temperature, move_begin=seg * per,
                           trace_every=fx.trace_every)
        bits.append(np.unpackbits(b, bitorder="little")[:per])
        trace.append(tr)
        cks.append(upd.get_cf_array())
    ref_bits = np.unpackbits(fx.accept_bits, bitorder="little")[:fx.n_moves]
    assert np.array_equal(np.concatenate(bits), ref_bits)
    assert np.array_equal(np.concatenate(trace), fx.e_trace)          # bit-identical energies
    assert np.array_equal(np.array(cks), fx.cf_checkpoints)           # bit-identical CFs
    assert np.array_equal(upd.get_occupation(), fx.occ_final)


def test_oracle_replays_reference_api_log():
    meta, bc = load_api_fixture()
    names = meta["names"]
    eci = seeded_ecis(bc, meta["eci_seed"], meta["eci_sigma"])
    occ = random_occupation(bc.n_sites, 3, meta["occ_seed"])
    upd = OracleUpdater(OracleModel(bc, names), occ, brute_force_cf(bc, occ, names), eci)
    for entry in meta["log"]:
        op, val = entry["op"], entry["value"]
        if op == "update_cf":
            upd.update_cf(tuple(val))
        elif op == "calculate":
            assert upd.calculate([tuple(c) for c in val[:-1]]) == val[-1]
        elif op == "undo_changes":
            upd.undo_changes()
        elif op == "clear_history":
            upd.clear_history()
        elif op == "set_ecis":
            upd.set_ecis(val)
        assert list(upd.get_cf_array()) == entry["cf"], op
        assert upd.get_energy() == entry["energy"], op
        assert upd.get_symbols() == entry["symbols"], op
        assert list(upd.get_singlets()) == entry["singlets"], op
    cur = upd.get_symbols()
    sp = bc.species
    with pytest.raises(RuntimeError, match="background"):
        upd.calculate([(5, cur[5], sp[(sp.index(cur[5]) + 1) % 3])])
    assert meta["errors"]["background"][0] == "RuntimeError"


@pytest.mark.parametrize("lattice,symmetrize", [("fcc", False), ("fcc", True), ("bcc", True), ("sc", True)])
def test_incremental_cf_equals_bruteforce(lattice, symmetrize):
    s = CubicSettings(5, ["Al", "Mg", "Si"], {2: 3, 3: 2, 4: 1}, symmetrize_decorations=symmetrize, lattice=lattice)
    names = s.cf_names()
    eci = seeded_ecis(s, 4, 0.01)
    occ = random_occupation(s.n_sites, 3, 8)
    upd = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    rng = np.random.Generator(np.random.Philox(12))
    cur = occ.copy()
    for step in range(40):
        i, j = (int(v) for v in rng.integers(0, s.n_sites, 2))
        if step % 2 and cur[i] != cur[j]:
            chs = [(i, int(cur[i]), int(cur[j])), (j, int(cur[j]), int(cur[i]))]
        else:
            chs = [(i, int(cur[i]), int((cur[i] + 1 + step % 2) % 3))]
        e = upd.calculate(chs)
        upd.clear_history()
        cur = upd.get_occupation()
        ref = brute_force_cf(s, cur, names)
        ref_arr = np.array([ref[n] for n in names])
        assert np.max(np.abs(upd.get_cf_array() - ref_arr)) < 1e-12
        assert abs(e - s.n_sites * float(np.dot([eci[n] for n in names], ref_arr))) < 1e-10


def _ising_delta(s, eci, occ, i):
    """Energy change of flipping site i in a binary +-1 system, from the
    geometric neighbour lists (independent of cluster_info)."""
    spin = np.where(occ == 0, 1.0, -1.0)  # phi(A)=+1, phi(B)=-1
    pos = s.positions
    L = np.array(s.dims)
    delta = 0.0
    cf_names = s.cf_names()
    if "c1_0" in cf_names:
        delta += eci["c1_0"] * (-2.0 * spin[i])
    for orb in s.orbits:
        name = orb.name + "_" + "0" * orb.size
        if name not in eci:
            continue
        acc = 0.0
        for cl in orb.offsets:  # clusters containing the origin, origin = site i
            prod = 1.0
            for v in cl[1:]:
                cell = s._to_cell(pos[i] + v)
                prod *= spin[int(s.site_index(cell % L))]
            acc += prod
        # E = N * J * cf, cf = (1/(M N)) * n * sum_distinct ; flipping s_i changes
        # each of the M clusters containing i by -2 s_i * (rest)
        delta += eci[name] * orb.size * (-2.0 * spin[i]) * acc / orb.multiplicity
    return delta


@pytest.mark.parametrize("lattice", ["fcc", "bcc", "sc"])
def test_binary_energy_change_equals_ising_local_field(lattice):
    s = CubicSettings(6, ["A", "B"], {2: 4, 3: 2}, lattice=lattice)
    names = s.cf_names()
    eci = seeded_ecis(s, 19, 0.02)
    occ = random_occupation(s.n_sites, 2, 23)
    upd = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    e0 = upd.get_energy()
    rng = np.random.Generator(np.random.Philox(2))
    for i in (int(v) for v in rng.integers(0, s.n_sites, 25)):
        occ_now = upd.get_occupation()
        closed = _ising_delta(s, eci, occ_now, i)
        e1 = upd.calculate([(i, int(occ_now[i]), 1 - int(occ_now[i]))])
        assert abs((e1 - e0) - closed) < 1e-12
        upd.undo_changes()
        assert upd.get_energy() == e0


@pytest.mark.parametrize("lattice,counts", [("fcc", (12, 6, 24, 12)), ("bcc", (8, 6, 12, 24)),
                                            ("sc", (6, 12, 8, 6))])
def test_pair_orbits_are_the_neighbour_shells(lattice, counts):
    """Pair orbits of each cubic lattice are its neighbour shells: the
    members of site 0 are exactly the sites at that minimum-image distance
    (from Cartesian positions), with the textbook coordination numbers."""
    s = CubicSettings(7, ["A", "B"], {2: 4}, lattice=lattice)
    pos = s.positions.astype(float)  # a/2 units
    box = np.array(s.dims)[:, None] * s.prim.astype(float)  # supercell vectors (rows)
    frac = (pos - pos[0]) @ np.linalg.inv(box)
    frac -= np.round(frac)
    d2 = np.round(((frac @ box) ** 2).sum(axis=1)).astype(int)
    shells = sorted(set(d2.tolist()) - {0})[:4]
    pairs = sorted((o for o in s.orbits if o.size == 2), key=lambda o: o.d2)
    assert [o.multiplicity for o in pairs] == list(counts)
    for o, r2 in zip(pairs, shells):
        assert o.d2 == r2
        members = sorted(m[0] for m in s.cluster_info[0][o.name]["indices"])
        assert members == sorted(np.nonzero(d2 == r2)[0].tolist())


def test_exhaustive_tiny_cell_every_state_every_flip():
    """2x2x3 cell with nearest-neighbour pairs and triplets: for all 4096
    states and every site, the incremental change equals the difference of
    from-scratch correlation functions."""
    s = FCCSettings((2, 2, 3), ["A", "B"], {2: 1, 3: 1})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    N = s.n_sites
    states = np.array(list(itertools.product([0, 1], repeat=N)), dtype=np.int8)
    model = OracleModel(s, names)
    rng = np.random.Generator(np.random.Philox(3))
    for occ in states[rng.permutation(len(states))[:300]]:
        cf_occ = brute_force_cf(s, occ, names)
        upd = OracleUpdater(model, occ, cf_occ, eci)
        for i in range(N):
            upd.update_cf((i, int(occ[i]), 1 - int(occ[i])))
            new = occ.copy()
            new[i] = 1 - new[i]
            ref = brute_force_cf(s, new, names)
            got = upd.get_cf()
            assert max(abs(got[n] - ref[n]) for n in names) < 1e-13
            upd.undo_changes()
            assert list(upd.get_occupation()) == list(occ)


def test_undo_restores_bit_identical_state_and_swaps_conserve_composition():
    fx = ChainFixture("ternary_canonical.npz")
    upd = OracleUpdater(OracleModel(fx.settings, fx.names), fx.occ0, fx.cf0, fx.eci)
    before_cf = upd.get_cf_array().copy()
    before_occ = upd.get_occupation().copy()
    e_before = upd.get_energy()
    i = 3
    j = next(k for k in range(fx.settings.n_sites) if before_occ[k] != before_occ[i])
    upd.calculate([(i, int(before_occ[i]), int(before_occ[j])), (j, int(before_occ[j]), int(before_occ[i]))])
    # the initial state occupies one slot (cpp/src/cf_history_tracker.cpp:46-56)
    assert upd.history_size() == 3
    upd.undo_changes()
    upd.update_cf((i, int(before_occ[i]), int((before_occ[i] + 1) % 3)))
    upd.update_cf((j, int(before_occ[j]), int((before_occ[j] + 2) % 3)))
    upd.undo_changes()
    assert np.array_equal(upd.get_cf_array(), before_cf)
    assert np.array_equal(upd.get_occupation(), before_occ)
    assert upd.get_energy() == e_before
    counts = np.bincount(before_occ, minlength=3)
    upd.run(0, 5, 1, 2000, 600.0)
    assert np.array_equal(np.bincount(upd.get_occupation(), minlength=3), counts)


def test_detailed_balance_tiny_cell_canonical():
    """Long canonical run on a 2x2x2 cell (4+4 of two species): visit
    frequencies of the 70 states follow the Boltzmann distribution."""
    s = FCCSettings((2, 2, 2), ["A", "B"], {2: 1})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    eci[[n for n in names if n.startswith("c2")][0]] = 0.004
    T = 300.0
    states = [np.array(c, dtype=np.int8) for c in itertools.product([0, 1], repeat=8) if sum(c) == 4]
    energies = []
    for st in states:
        cf = brute_force_cf(s, st, names)
        energies.append(s.n_sites * sum(eci[n] * cf[n] for n in names))
    energies = np.array(energies)
    w = np.exp(-(energies - energies.min()) / (KB_EV * T))
    p = w / w.sum()
    model = OracleModel(s, names)
    occ = states[0]
    upd = OracleUpdater(model, occ, brute_force_cf(s, occ, names), eci)
    index = {tuple(st): k for k, st in enumerate(states)}
    counts = np.zeros(len(states))
    for blk in range(4000):
        upd.run(0, 77, 0, 25, T, move_begin=blk * 25)
        counts[index[tuple(upd.get_occupation())]] += 1
    exp = p * counts.sum()
    chi2 = float(np.sum((counts - exp) ** 2 / exp))
    # 69 degrees of freedom; samples 25 moves apart are nearly independent
    assert chi2 < 130, chi2


def test_sgc_singlet_equals_composition():
    """For the binary +-1 basis the singlet is (n_A - n_B) / N."""
    s = FCCSettings(6, ["A", "B"], {2: 2})
    names = s.cf_names()
    eci = seeded_ecis(s, 1, 0.02)
    occ = random_occupation(s.n_sites, 2, 4)
    upd = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    upd.run(1, 3, 0, 3000, 900.0)
    o = upd.get_occupation()
    assert abs(upd.get_singlets()[0] - (np.sum(o == 0) - np.sum(o == 1)) / s.n_sites) < 1e-13


def test_metropolis_acceptance_rule_matches_reference_formula():
    """The oracle's accepted energies follow the reference rule: every
    accepted uphill move satisfies u <= exp(-dE / (kB T))."""
    fx = ChainFixture("binary_sgc_chain1.npz")
    upd = OracleUpdater(OracleModel(fx.settings, fx.names), fx.occ0, fx.cf0, fx.eci_run)
    e_prev = upd.get_energy()
    for t in range(300):
        bits, tr, _ = upd.run(fx.kind, fx.seed, fx.chain_uid, 1, fx.temperature, move_begin=t, trace_every=1)
        w = philox4x64((t, 0, 0, 0), (fx.seed, fx.chain_uid))
        u = float(w[3] >> 11) * 2.0 ** -53
        de = tr[0] - e_prev
        if bits[0] & 1 and de > 0:
            assert u <= math.exp(-de / (fx.temperature * KB_EV))
        e_prev = tr[0]
