"""GPU parity against the reference fixtures and the oracle.

Acceptance sequences must be bit-identical; energies within 1e-10 eV/site and
correlation functions within 1e-10 of the reference's (the tolerance of the
target: |E_gpu - E_ref| / N <= 1e-10 eV/site).
"""
import numpy as np
import pytest

from golden_io import CHAIN_FIXTURES, CRYSTAL_FIXTURES, ChainFixture, load_api_fixture

E_TOL = 1e-10  # eV per site
CF_TOL = 1e-10

pytestmark = pytest.mark.gpu


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


def _batch_for(fx, settings=None, path=0):
    from cemc_amd.batch import ChainBatch
    from cemc_amd.model import CEModel
    model = CEModel(settings if settings is not None else fx.settings, fx.eci)
    chain = dict(kind=fx.kind, temperature=fx.temperature, occupation=fx.occ0, seed=fx.seed,
                 uid=fx.chain_uid, chem_pot=fx.meta["mu"] or None)
    batch = ChainBatch([model], [chain])
    if path:
        batch.set_energy_path(path)
    return model, batch


@pytest.mark.parametrize("path", ["evaluate", "fields"])
@pytest.mark.parametrize("fname", CHAIN_FIXTURES + CRYSTAL_FIXTURES)
def test_batch_reproduces_reference_chain(gpu, fname, path):
    """Both energy paths: tables summed per trial move, and per-site fields
    updated on accepted moves (periodic fixtures; the others evaluate)."""
    from cemc_amd.batch import PATH_EVALUATE, PATH_FIELDS
    fx = ChainFixture(fname)
    print("fixture", fname, "moves", fx.n_moves, "path", path, flush=True)
    model, batch = _batch_for(fx, path=PATH_FIELDS if path == "fields" else PATH_EVALUATE)
    if path == "fields" and model.periodic and model.n_subclusters <= 512:
        assert batch.energy_path(0) == PATH_FIELDS
    n = fx.n_sites
    bits = []
    for seg in range(fx.n_moves // fx.cf_every):
        for sub in range(fx.cf_every // fx.trace_every):
            log = batch.run_logged(fx.trace_every)
            bits.append(_bits(log[0], fx.trace_every))
            idx = seg * (fx.cf_every // fx.trace_every) + sub
            e = batch.energies()[0]
            assert abs(e - fx.e_trace[idx]) / n <= E_TOL, (idx, e, fx.e_trace[idx])
        cf, e_cf = batch.cf(0)
        assert np.max(np.abs(cf - fx.cf_checkpoints[seg])) <= CF_TOL
    got = np.concatenate(bits)
    ref = _bits(fx.accept_bits, fx.n_moves)
    mism = np.nonzero(got != ref)[0]
    assert mism.size == 0, "first mismatching move %d" % (mism[0] if mism.size else -1)
    assert np.array_equal(batch.occupation(0), fx.occ_final)
    st = batch.stats()
    assert st["accepted"][0] == fx.n_accept
    assert st["moves"][0] == fx.n_moves


def test_periodic_and_table_paths_agree(gpu):
    """The closed-form neighbour path reproduces the translation-table path."""
    fx = ChainFixture("binary_canonical_plain.npz")
    model_p, batch_p = _batch_for(fx)
    assert model_p.periodic

    class NoDims:
        pass

    plain = NoDims()
    for key, val in vars(fx.settings).items():
        if key not in ("dims", "cells"):
            setattr(plain, key, val)
    model_t, batch_t = _batch_for(fx, plain)
    assert not model_t.periodic
    a = batch_p.run_logged(20000)
    b = batch_t.run_logged(20000)
    assert np.array_equal(a, b)
    assert np.array_equal(batch_p.occupation(0), batch_t.occupation(0))
    assert abs(batch_p.energies()[0] - batch_t.energies()[0]) <= 1e-12 * fx.n_sites


def test_initial_cf_matches_bruteforce(gpu):
    """GPU correlation functions from scratch vs the oracle's brute force, with
    symmetrised decorations (equivalent sites) on a ternary and a quaternary cell."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    from oracle.brute_force import brute_force_cf
    for species, shells, L in ((["A", "B", "C"], {2: 4, 3: 3, 4: 2}, 6), (["A", "B", "C", "D"], {2: 3, 3: 3, 4: 3}, 5)):
        s = FCCSettings(L, species, shells)
        eci = seeded_ecis(s, 5, 0.01)
        occ = random_occupation(s.n_sites, len(species), 17)
        model = CEModel(s, eci)
        batch = ChainBatch([model], [dict(kind=0, temperature=1000.0, occupation=occ, seed=1, uid=0)])
        cf, e = batch.cf(0)
        ref = brute_force_cf(s, occ, model.names)
        ref_arr = np.array([ref[n] for n in model.names])
        assert np.max(np.abs(cf - ref_arr)) <= 1e-12
        e_ref = s.n_sites * float(np.dot(model.eci, ref_arr))
        assert abs(e - e_ref) / s.n_sites <= 1e-12
        assert abs(batch.energies()[0] - e) <= 1e-12 * s.n_sites


def test_updater_replays_reference_api_log(gpu):
    """PyCEUpdater on the GPU against the reference's recorded call sequence."""
    from cemc_amd.lattice import SimpleAtoms, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    from oracle.brute_force import brute_force_cf
    meta, bc = load_api_fixture()
    names = meta["names"]
    eci = seeded_ecis(bc, meta["eci_seed"], meta["eci_sigma"])
    occ = random_occupation(bc.n_sites, 3, meta["occ_seed"])
    cf0 = brute_force_cf(bc, occ, names)
    atoms = SimpleAtoms([bc.species[s] for s in occ])
    upd = PyCEUpdater(atoms, bc, cf0, eci)
    n = bc.n_sites

    def check(entry):
        cf = upd.get_cf()
        assert max(abs(cf[k] - v) for k, v in zip(names, entry["cf"])) <= CF_TOL, entry["op"]
        assert abs(upd.get_energy() - entry["energy"]) / n <= E_TOL, entry["op"]
        assert upd.get_symbols() == entry["symbols"], entry["op"]
        assert [a.symbol for a in atoms] == entry["symbols"], entry["op"]
        assert np.allclose(upd.get_singlets(), entry["singlets"], atol=CF_TOL, rtol=0)

    for entry in meta["log"]:
        op, val = entry["op"], entry["value"]
        if op == "update_cf":
            upd.update_cf(tuple(val))
        elif op == "calculate":
            e = upd.calculate([tuple(c) for c in val[:-1]])
            assert abs(e - val[-1]) / n <= E_TOL
        elif op == "undo_changes":
            upd.undo_changes()
        elif op == "clear_history":
            upd.clear_history()
        elif op == "set_ecis":
            upd.set_ecis(val)
        check(entry)
    kind, _ = meta["errors"]["background"]
    cur = upd.get_symbols()
    sp = bc.species
    with pytest.raises(RuntimeError):
        upd.calculate([(5, cur[5], sp[(sp.index(cur[5]) + 1) % 3])])
    assert kind == "RuntimeError"


def test_updater_matches_oracle_with_equivalent_sites(gpu):
    """Random updater call sequences on a ternary cell whose clusters have
    equivalent sites (symmetrised decorations): GPU vs the C oracle."""
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    s = FCCSettings(5, ["Al", "Mg", "Si"], {2: 3, 3: 2, 4: 1})
    names = s.cf_names()
    eci = seeded_ecis(s, 9, 0.01)
    occ = random_occupation(s.n_sites, 3, 21)
    cf0 = brute_force_cf(s, occ, names)
    upd = PyCEUpdater(SimpleAtoms([s.species[v] for v in occ]), s, cf0, eci)
    ora = OracleUpdater(OracleModel(s, names), occ, cf0, eci)
    rng = np.random.Generator(np.random.Philox(3))
    cur = list(occ)
    for step in range(60):
        i = int(rng.integers(0, s.n_sites))
        j = int(rng.integers(0, s.n_sites))
        if cur[i] != cur[j] and step % 3 == 0:
            chs = [(i, s.species[cur[i]], s.species[cur[j]]), (j, s.species[cur[j]], s.species[cur[i]])]
        else:
            chs = [(i, s.species[cur[i]], s.species[(cur[i] + 1 + step % 2) % 3])]
        e_gpu = upd.calculate(chs)
        e_ora = ora.calculate(chs)
        assert abs(e_gpu - e_ora) / s.n_sites <= 1e-12
        if step % 4 == 1:
            upd.undo_changes()
            ora.undo_changes()
        else:
            upd.clear_history()
            ora.clear_history()
        cur = [s.species.index(v) for v in upd.get_symbols()]
        assert cur == list(ora.get_occupation())
        assert np.max(np.abs(upd.get_cf_array() - ora.get_cf_array())) <= 1e-12


def _allowed_species(s):
    """Species ids each site may hold (its sublattice's elements; none for
    background sites)."""
    sp = list(s.species)
    bkg = set(int(i) for i in s.background_indices)
    sub = getattr(s, "sublattice_of_site", None)
    out = []
    for i in range(s.n_sites):
        if i in bkg:
            out.append([])
        elif sub is None:
            out.append(list(range(len(sp))))
        else:
            out.append([sp.index(e) for e in s.basis_elements[int(sub[i])]])
    return out


def _replay_random_calls(s, occ, steps, seed, cf0=None, eci=None):
    """Random updater calls on the GPU and on the oracle: single flips, swaps
    of unlike sites of one sublattice, and twelve flips in one call (more
    changes than travel in the kernel arguments); undo or clear after each.
    Energies, correlation functions and lattices are compared after every
    call."""
    from cemc_amd.lattice import SimpleAtoms, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    if eci is None:
        eci = seeded_ecis(s, 4, 0.01)
    names = sorted(eci)
    sp = list(s.species)
    atoms = SimpleAtoms([sp[v] for v in occ])
    upd = PyCEUpdater(atoms, s, cf0 if cf0 is None else dict(cf0), eci)
    if cf0 is None:
        cf0 = brute_force_cf(s, occ, names) if s.n_sites <= 20000 else upd.get_cf()
    ora = OracleUpdater(OracleModel(s, names), occ, cf0, eci)
    allowed = _allowed_species(s)
    movable = [i for i in range(s.n_sites) if len(allowed[i]) > 1]
    rng = np.random.Generator(np.random.Philox(seed))
    cur = [int(v) for v in occ]

    def flip(state):
        i = movable[int(rng.integers(len(movable)))]
        opts = [v for v in allowed[i] if v != state[i]]
        new = opts[int(rng.integers(len(opts)))]
        ch = (i, sp[state[i]], sp[new])
        state[i] = new
        return ch

    n_long = 0
    for step in range(steps):
        tmp = list(cur)
        if step % 3 == 0:
            chs = [flip(tmp) for _ in range(12)]
            n_long += 1
        elif step % 3 == 1:
            i = movable[int(rng.integers(len(movable)))]
            js = [j for j in movable if allowed[j] == allowed[i] and cur[j] != cur[i]]
            if not js:
                chs = [flip(tmp)]
            else:
                j = js[int(rng.integers(len(js)))]
                chs = [(i, sp[cur[i]], sp[cur[j]]), (j, sp[cur[j]], sp[cur[i]])]
        else:
            chs = [flip(tmp)]
        e_gpu = upd.calculate(chs)
        e_ora = ora.calculate(chs)
        assert abs(e_gpu - e_ora) / s.n_sites <= 1e-12, step
        if step % 4 == 1:
            upd.undo_changes()
            ora.undo_changes()
        else:
            upd.clear_history()
            ora.clear_history()
        cur = [sp.index(v) for v in upd.get_symbols()]
        assert cur == [int(v) for v in ora.get_occupation()], step
        assert [a.symbol for a in atoms] == upd.get_symbols(), step
        assert np.max(np.abs(upd.get_cf_array() - ora.get_cf_array())) <= 1e-12, step
        assert abs(upd.get_energy() - ora.get_energy()) / s.n_sites <= 1e-12, step
    assert n_long > 0
    upd.close()


def _crystal(kind):
    from cemc_amd.crystal import HCPSettings, Spacegroup217Settings
    from cemc_amd.lattice import CrystalSettings
    if kind == "rocksalt":
        return CrystalSettings(3, [["Na", "K"], ["Cl", "Br"]], {2: 3, 3: 2}, "rocksalt")
    if kind == "rocksalt_background":
        return CrystalSettings(3, [["Na", "K", "Li"], ["Cl"]], {2: 3, 3: 2}, "rocksalt")
    if kind == "hcp_ternary":
        return HCPSettings((3, 3, 2), ["Al", "Mg", "Si"], {2: 1.45, 3: 1.05, 4: 1.05})
    wyckoff = [(0, 0, 0), (0.324, 0.324, 0.324), (0.3582, 0.3582, 0.0393), (0.0954, 0.0954, 0.2725)]
    return Spacegroup217Settings(1, 10.553, wyckoff, [["Al", "Mg"]] * 4, {2: 4.1, 3: 4.1})


@pytest.mark.parametrize("kind", ["rocksalt", "rocksalt_background", "hcp_ternary", "sg217_binary"])
def test_updater_matches_oracle_on_crystals_and_long_change_lists(gpu, kind):
    """Crystals whose lattices come from translation tables: two translation
    groups (rocksalt, with and without a background sublattice), one group
    of two sites (HCP) and 29 groups (I-43m), so consecutive changes of a
    call switch groups; change lists of twelve flips are read from device
    memory."""
    from cemc_amd.lattice import random_crystal_occupation
    s = _crystal(kind)
    print("updater on", kind, s.n_sites, "sites", flush=True)
    _replay_random_calls(s, random_crystal_occupation(s, 5), 45, 11)


def test_undo_after_more_changes_than_the_history_holds(gpu):
    """1005 single-site changes without a clear: the history keeps the last
    999 (the ring of 1000 slots), an undo restores the state after the
    sixth change on the device, in the atoms object and in the oracle."""
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    from oracle.brute_force import brute_force_cf
    from oracle.oracle import OracleModel, OracleUpdater
    s = FCCSettings(4, ["Al", "Mg"], {2: 2})
    names = s.cf_names()
    eci = seeded_ecis(s, 2, 0.01)
    occ = random_occupation(s.n_sites, 2, 4)
    atoms = SimpleAtoms([s.species[v] for v in occ])
    cf0 = brute_force_cf(s, occ, names)
    upd = PyCEUpdater(atoms, s, cf0, eci)
    ora = OracleUpdater(OracleModel(s, names), occ, cf0, eci)
    cur = [int(v) for v in occ]
    rng = np.random.Generator(np.random.Philox(3))
    for _ in range(1005):
        i = int(rng.integers(s.n_sites))
        ch = (i, s.species[cur[i]], s.species[1 - cur[i]])
        cur[i] = 1 - cur[i]
        upd.update_cf(ch)
        ora.update_cf(ch)
    assert upd.history_size() == ora.history_size()
    upd.undo_changes()
    ora.undo_changes()
    assert [int(v) for v in upd.get_occupation()] == [int(v) for v in ora.get_occupation()]
    assert [a.symbol for a in atoms] == upd.get_symbols()
    assert np.max(np.abs(upd.get_cf_array() - ora.get_cf_array())) <= 1e-12
    upd.close()


def test_updater_with_singlet_correlation_functions_only(gpu):
    """ECIs on c0 and the singlets only: a change touches no cluster
    correlation function (no plan items), in calls of up to twelve
    changes."""
    from cemc_amd.lattice import FCCSettings, random_occupation
    s = FCCSettings(4, ["Al", "Mg", "Si"], {2: 1})
    print("updater with singlets only", flush=True)
    _replay_random_calls(s, random_occupation(s.n_sites, 3, 7), 30, 13,
                         eci={"c0": 0.05, "c1_0": 0.3, "c1_1": -0.2})


def test_updater_on_a_lattice_larger_than_lds(gpu):
    """FCC 54^3 (157464 sites): the occupation does not fit in LDS beside
    the rest, so the updater reads it from global memory, with full
    barriers between the changes of a call."""
    from cemc_amd.lattice import FCCSettings, random_occupation
    s = FCCSettings(54, ["A", "B"], {2: 1, 3: 1})
    print("updater on FCC 54^3", flush=True)
    _replay_random_calls(s, random_occupation(s.n_sites, 2, 3), 30, 12)


@pytest.mark.parametrize("case", ["fcc_binary", "rocksalt", "fcc54_occupation_in_hbm"])
def test_resident_updater_equals_one_launch_per_call(gpu, case):
    """The resident kernel (calls as commands taken from mapped host memory
    by a workgroup that stays on the GPU) and one kernel launch per call give
    identical energies, correlation functions and lattices over the same
    random calls: one, two, seven, eight and twelve changes per call (seven
    fill a command, or follow restores that no longer fit beside them; eight
    and twelve go to a launch), update_cf without a result, undos of up to 40
    changes (more than 30 go to a launch), get_energy and get_cf after undos, new ECIs and a new
    occupation midway, and pauses longer than the resident kernel's 0.5 ms
    idle time, after which it is launched again.  The resident path serves
    an updater while it is the only live one with the path switched on, so
    updaters left over from earlier tests are collected first."""
    import gc
    import time
    gc.collect()
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_crystal_occupation, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    if case == "fcc_binary":
        s, steps = FCCSettings(6, ["A", "B"], {2: 3, 3: 2}), 240
    elif case == "rocksalt":
        s, steps = _crystal("rocksalt"), 160
    else:
        s, steps = FCCSettings(54, ["A", "B"], {2: 1, 3: 1}), 60
    occ = random_occupation(s.n_sites, 2, 5) if case != "rocksalt" else random_crystal_occupation(s, 5)
    eci = seeded_ecis(s, 8, 0.01)
    sp = list(s.species)
    ups = []
    for resident in (True, False):
        u = PyCEUpdater(SimpleAtoms([sp[v] for v in occ]), s, None, dict(eci))
        u.set_resident(resident)
        ups.append(u)
    allowed = _allowed_species(s)
    movable = [i for i in range(s.n_sites) if len(allowed[i]) > 1]
    rng = np.random.Generator(np.random.Philox(21))
    cur = [int(v) for v in occ]
    print("resident vs launch per call on", case, s.n_sites, "sites", flush=True)

    def flips(k):
        state = list(cur)
        out = []
        for _ in range(k):
            i = movable[int(rng.integers(len(movable)))]
            opts = [v for v in allowed[i] if v != state[i]]
            new = opts[int(rng.integers(len(opts)))]
            out.append((i, sp[state[i]], sp[new]))
            state[i] = new
        return out

    def same(step):
        e = [u.get_energy() for u in ups]
        assert e[0] == e[1], step
        assert np.array_equal(ups[0].get_cf_array(), ups[1].get_cf_array()), step
        assert np.array_equal(ups[0].get_occupation(), ups[1].get_occupation()), step

    for step in range(steps):
        kind = step % 8
        if kind in (0, 1, 2, 3):
            chs = flips((1, 2, 12, 7)[kind])
            e = [u.calculate(chs) for u in ups]
            assert e[0] == e[1], step
        elif kind == 4:
            for ch in flips(int(rng.integers(20, 41))):
                for u in ups:
                    u.update_cf(ch)
        elif kind == 5:
            for ch in flips(5):
                for u in ups:
                    u.update_cf(ch)
        elif kind == 6:
            time.sleep(0.005)
        else:
            for ch in flips(3):
                e = [u.calculate([ch]) for u in ups]
                assert e[0] == e[1], step
            cur = [int(v) for v in ups[0].get_occupation()]
            chs = flips(8)
            e = [u.calculate(chs) for u in ups]
            assert e[0] == e[1], step
        if step % 3 == 0:
            for u in ups:
                u.undo_changes()
        else:
            for u in ups:
                u.clear_history()
        same(step)
        cur = [int(v) for v in ups[0].get_occupation()]
        if step == steps // 2:
            new_eci = {k: v * 1.5 for k, v in seeded_ecis(s, 9, 0.01).items()}
            for u in ups:
                u.set_ecis(new_eci)
            same(step)
        if step == 3 * steps // 4:
            occ2 = random_occupation(s.n_sites, 2, 6) if case != "rocksalt" else random_crystal_occupation(s, 6)
            for u in ups:
                u.set_occupation(occ2)
            cur = [int(v) for v in occ2]
            same(step)
    for u in ups:
        u.close()


def test_resident_updater_hands_over_when_a_second_updater_appears(gpu):
    """An updater on the resident path while a second updater with the path
    on is created (both then launch per call, the first after running the
    restores that were waiting for its next command) and closed again (the
    first returns to the resident path) gives the same energies, correlation
    functions and lattice as an updater with one launch per call throughout."""
    import gc
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    gc.collect()
    s = FCCSettings(6, ["A", "B"], {2: 3, 3: 2})
    eci = seeded_ecis(s, 8, 0.01)
    occ = random_occupation(s.n_sites, 2, 5)
    sp = list(s.species)
    a = PyCEUpdater(SimpleAtoms([sp[v] for v in occ]), s, None, dict(eci))
    ref = PyCEUpdater(SimpleAtoms([sp[v] for v in occ]), s, None, dict(eci))
    ref.set_resident(False)
    other = None
    rng = np.random.Generator(np.random.Philox(33))
    print("resident updater handing over", flush=True)
    for step in range(90):
        cur = [int(v) for v in a.get_occupation()]
        i, j = (int(v) for v in rng.choice(s.n_sites, 2, replace=False))
        if cur[i] != cur[j]:
            chs = [(i, sp[cur[i]], sp[cur[j]]), (j, sp[cur[j]], sp[cur[i]])]
        else:
            chs = [(i, sp[cur[i]], sp[1 - cur[i]])]
        assert a.calculate(chs) == ref.calculate(chs), step
        if step % 3 != 2:
            a.undo_changes()
            ref.undo_changes()
        else:
            a.clear_history()
            ref.clear_history()
        if step == 30:
            other = PyCEUpdater(SimpleAtoms([sp[v] for v in occ]), s, None, dict(eci))
        if step == 60:
            other.close()
        if step % 10 == 0:
            assert a.get_energy() == ref.get_energy(), step
            assert np.array_equal(a.get_cf_array(), ref.get_cf_array()), step
        assert np.array_equal(a.get_occupation(), ref.get_occupation()), step
    a.close()
    ref.close()


@pytest.mark.parametrize("fname,n", [("binary_canonical_plain.npz", 3000), ("ternary_sgc.npz", 3000),
                                     ("quaternary_canonical.npz", 400)])
def test_dropin_updater_loop_reproduces_reference_chain(gpu, fname, n):
    """The drop-in path: the per-move loop a caller of PyCEUpdater runs (one
    calculate, then clear_history or undo_changes, per move, as the
    reference's Montecarlo._mc_step does through CE) reproduces the
    reference's chain move for move on the fed stream."""
    from cemc_amd.dropin import run_updater_chain
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.updater import PyCEUpdater
    fx = ChainFixture(fname)
    species = sorted(fx.settings.unique_elements)
    atoms = SimpleAtoms([species[s] for s in fx.occ0])
    upd = PyCEUpdater(atoms, fx.settings, dict(fx.cf0), dict(fx.eci_run))
    print("drop-in loop", fname, n, flush=True)
    bits, e_trace, _, final, n_acc = run_updater_chain(upd, atoms, species, fx.kind, fx.seed, fx.chain_uid, n,
                                                       fx.temperature, trace_every=fx.trace_every)
    assert np.array_equal(_bits(bits, n), _bits(fx.accept_bits, n))
    ref_e = fx.e_trace[:len(e_trace)]
    assert np.max(np.abs(e_trace - ref_e)) / fx.n_sites <= E_TOL
    if n == fx.n_moves:
        assert np.array_equal(final, fx.occ_final)
        assert n_acc == fx.n_accept
    # the atoms object follows the lattice, as the reference's updater keeps it
    assert [a.symbol for a in atoms] == upd.get_symbols()


def test_updater_singlets_vibrations_and_threads(gpu):
    """get_singlets, add_linear_vib_correction / vib_energy and
    set_num_threads through the C ABI (ce_updater.cpp:655-698,
    linear_vib_correction.cpp:8-16)."""
    import ctypes

    from cemc_amd import _capi
    from cemc_amd.lattice import FCCSettings, SimpleAtoms, random_occupation, seeded_ecis
    from cemc_amd.updater import PyCEUpdater
    from oracle.brute_force import brute_force_cf
    s = FCCSettings(5, ["Al", "Mg", "Si"], {2: 2, 3: 1})
    names = s.cf_names()
    eci = seeded_ecis(s, 3, 0.02)
    occ = random_occupation(s.n_sites, 3, 8)
    atoms = SimpleAtoms([s.species[v] for v in occ])
    upd = PyCEUpdater(atoms, s, None, dict(eci))
    upd.calculate([(0, atoms[0].symbol, "Si" if atoms[0].symbol != "Si" else "Al")])
    upd.clear_history()
    cf = upd.get_cf()
    sing = upd.get_singlets()
    assert np.array_equal(sing, [cf["c1_0"], cf["c1_1"]])
    assert np.allclose(sing, [brute_force_cf(s, upd.get_occupation(), names)[k] for k in ("c1_0", "c1_1")],
                       atol=1e-12)
    lib = _capi.load()
    small = np.zeros(1)
    rc = lib.cemc_updater_get_singlets(upd.handle, small.ctypes.data_as(_capi._dp), 1, None)
    assert rc == _capi.CEMC_ERUNTIME
    assert "too small to hold all the singlets" in lib.cemc_last_error().decode()
    assert upd.vib_energy(800.0) == 0.0
    pair = [n for n in names if n.startswith("c2")][0]
    corr = {pair: 0.3, "c1_1": -0.2, "c0": 0.05}
    upd.add_linear_vib_correction(corr)
    kb = 8.6173303e-5
    want = 0.0
    for name in sorted(corr):
        want += cf[name] * corr[name] * kb * 800.0
    assert upd.vib_energy(800.0) == want
    with pytest.raises(ValueError):
        upd.add_linear_vib_correction({"c9_nonexistent": 1.0})
    upd.set_num_threads(4)
    with pytest.raises(ValueError):
        upd.set_num_threads(0)
    n = ctypes.c_int32(0)
    _capi.check(lib.cemc_updater_get_singlets(upd.handle, None, 0, ctypes.byref(n)))
    assert n.value == 2
