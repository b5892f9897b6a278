"""Wang-Landau schemes of WangLandauSampler::run beyond fixed f: adaptive
energy windows (cpp/src/adaptive_windows.cpp, wang_landau_sampler.cpp:575-580)
with their sub-bin counts (histogram.cpp:92-100), and the inverse-time
modification factor (wang_landau_sampler.cpp:344-367).

CPU, the oracle's restatement pinned on a 2x2x4 binary cell whose 12870
states at fixed composition are enumerated (76 energy levels, 32 occupied
bins of 50):
* adaptive windows cut the range several times and still converge to the
  exact ln g; the tolerance fails a flipped or missing shift at the cuts and
  a missing restore of the saved bins (each was planted in a copy of the
  oracle: errors 5.9e4, 927 and 4.3 against 0.84 here);
* one cut, followed on preset counts: position of the cut, the saved ln g
  above it, the new window, the walker moved to the recorded state nearest
  the window's middle (its lattice is that state's, which the enumeration
  confirms by energy);
* the reset after a flat window: ln g continuous at the cut with the saved
  values back above it, the walker back at the first cut's state;
* sub-bins: with ln g held at 0 every state is equally likely, so the
  sub-bin counts of each bin follow the number of states at each energy
  level in it, and only sub-bins holding a level are ever counted;
* inverse time: f halves until a flat histogram takes more steps than the
  previous one, then f t stays at its value at that flat (t = steps / bins),
  and the walker still converges toward the exact ln g.

GPU: the kernel reproduces the oracle walker exactly under every scheme.
"""
import itertools

import numpy as np
import pytest

from cemc_amd.lattice import FCCSettings
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater, OracleWL, wl_run

NB, W = 50, 0.0043
EMIN = -0.13003684210526315  # no energy level within 3 % of a bin edge
EMAX = EMIN + NB * W


@pytest.fixture(scope="module")
def cell():
    s = FCCSettings((2, 2, 4), ["A", "B"], {2: 3, 3: 1})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    c2 = [n for n in names if n.startswith("c2")]
    c3 = [n for n in names if n.startswith("c3")]
    eci[c2[0]], eci[c2[1]], eci[c2[2]], eci[c3[0]] = 0.01, -0.004, 0.0023, 0.003
    N = s.n_sites
    states = [np.array(c, dtype=np.int8) for c in itertools.product([0, 1], repeat=N) if sum(c) == N // 2]
    cfs = [brute_force_cf(s, st, names) for st in states]
    energies = np.array([N * sum(eci[n] * cf[n] for n in names) for cf in cfs])
    x = (energies - EMIN) * NB / (EMAX - EMIN)
    assert np.min(np.abs(x - np.round(x))) > 0.03
    g = np.bincount(np.trunc(x).astype(int), minlength=NB).astype(float)
    assert g.size == NB and (g > 0).sum() == 32
    return dict(s=s, names=names, eci=eci, states=states, energies=energies, g=g, model=OracleModel(s, names))


def _updater(c, k):
    occ = c["states"][k]
    return OracleUpdater(c["model"], occ, brute_force_cf(c["s"], occ, c["names"]), c["eci"])


def _walker(c, **kw):
    args = dict(f_min=1e-5, flatness=0.8, n_sites=c["s"].n_sites, n_cf=len(c["names"]))
    args.update(kw)
    f = args.pop("f", 1.0)
    return OracleWL(NB, EMIN, EMAX, f, **args)


def _ln_g_error(c, wl):
    g = c["g"]
    known = wl.known != 0
    assert np.array_equal(known, g > 0)
    d = (wl.logdos - np.log(np.where(g > 0, g, 1.0)))[known]
    return np.abs(d - d.mean()).max()


def _energy_of(c, occ):
    k = next(i for i, st in enumerate(c["states"]) if np.array_equal(st, occ))
    return c["energies"][k]


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_adaptive_windows_converge_to_exact_dos(cell, seed):
    upd = _updater(cell, 100 + seed)
    wl = _walker(cell, adaptive=True, min_width=3)
    used = []
    wl_run(upd, wl, seed, 1, 3000, 1000, counters=used)
    st = wl.state
    assert st.converged == 1 and st.n_flat == 17 and st.has_first == 1
    assert st.upper == NB and st.n_edges == 0
    # ln g spans ln(1440 / 4) = 5.9; the windows join at single bins, whose
    # noise the final stages do not remove (the reference marks the joining
    # as approximate, adaptive_windows.hpp:33)
    assert _ln_g_error(cell, wl) < 1.5
    # steps and moves back into a window use the counters in order
    assert used[0] == st.n_moves + st.n_walk


def test_one_cut_on_preset_counts(cell):
    """Record states with ln g fixed (f = 0) and no window flat (criteria
    2), then preset counts whose upper part is flat down to a low bin."""
    upd = _updater(cell, 7)
    wl = _walker(cell, f=0.0, f_min=-1.0, flatness=2.0, adaptive=True, min_width=3)
    wl_run(upd, wl, 4, 1, 20, 1000)
    assert wl.state.n_edges == 0 and wl.state.upper == NB
    known = wl.known != 0
    # a state is recorded at the first visit of every visited bin
    assert np.array_equal(wl.saved != 0, known) and np.all(cell["g"][known] > 0) and known.sum() > 20
    kb = np.flatnonzero(known)
    low = kb[len(kb) // 2]
    hist = np.where(known, 100000, 0).astype(np.int32)
    hist[low] = 10
    wl.hist[:] = hist
    wl.logdos[:] = np.arange(NB) * 0.01 + 1.0
    wl.state.flatness = 0.8
    lg_before = wl.logdos.copy()
    wl_run(upd, wl, 4, 1, 1, 1, move_begin=20000)
    st = wl.state
    # the scan starts at upper - n_overlap = 43 and stops at the low bin;
    # the window above it holds more than min_width known bins
    assert int(np.sum(known[low + 1:NB - 7])) > 3
    assert st.n_edges == 1 and wl.edges[0] == low + 1 and wl.edge_n[0] == 7
    assert np.array_equal(wl.edge_vals[:7], lg_before[low + 1:low + 8])
    assert st.upper == low + 8
    # the walker moved to the recorded state nearest the window's middle
    mid = st.upper // 2
    cand = [b for b in range(st.upper) if known[b]]
    want = min(cand, key=lambda b: (abs(b - mid), b < mid))
    assert st.cur_bin == want
    occ = upd.get_occupation()
    assert np.array_equal(occ, wl.snap_occ.reshape(NB + 1, -1)[want])
    e = _energy_of(cell, occ)
    assert int((e - EMIN) * NB / (EMAX - EMIN)) == want
    assert abs(upd.get_energy() - e) < 1e-12
    # the state at the cut is kept for the reset
    assert st.has_first == 1


def test_reset_joins_windows_continuously(cell):
    upd = _updater(cell, 9)
    wl = _walker(cell, f=0.0, f_min=-1.0, flatness=2.0, adaptive=True, min_width=3)
    wl_run(upd, wl, 5, 1, 20, 1000)
    known = wl.known != 0
    kb = np.flatnonzero(known)
    low = kb[len(kb) // 2]
    hist = np.where(known, 100000, 0).astype(np.int32)
    hist[low] = 10
    wl.hist[:] = hist
    lg0 = np.arange(NB) * 0.01 + 1.0
    wl.logdos[:] = lg0
    wl.state.flatness = 0.8
    wl_run(upd, wl, 5, 1, 1, 1, move_begin=20000)
    assert wl.state.n_edges == 1
    first_occ = wl.snap_occ.reshape(NB + 1, -1)[NB].copy()
    edge = int(wl.edges[0])
    # raise ln g inside the new window, then make it flat: the reset puts
    # the saved values back above the cut and shifts the window down
    lg1 = wl.logdos.copy()
    lg1[:wl.state.upper] += 3.0
    wl.logdos[:] = lg1
    wl.hist[:] = np.where(known, 1000, 0)
    wl_run(upd, wl, 5, 1, 1, 1, move_begin=30000)
    st = wl.state
    assert st.n_flat == 1 and st.upper == NB and st.n_edges == 0
    expect = lg1.copy()
    expect[edge + 1:edge + 7] = lg0[edge + 1:edge + 7]
    expect[:edge + 1] -= lg1[edge] - lg0[edge]
    # one visit of f = 0 changed nothing else
    assert np.allclose(wl.logdos, expect, rtol=0, atol=1e-12)
    assert np.array_equal(upd.get_occupation(), first_occ)


def test_sub_bins_follow_level_degeneracies(cell):
    """ln g = 0 throughout (f = 0, never flat): states are sampled
    uniformly; within each bin the sub-bin counts follow the number of
    states at each energy level."""
    upd = _updater(cell, 11)
    wl = _walker(cell, f=0.0, f_min=-1.0, flatness=2.0, adaptive=True, min_width=3)
    n = 200 * 1000
    wl_run(upd, wl, 6, 1, 200, 1000)
    sub = wl.sub_bins().astype(float)
    e = cell["energies"]
    x = (e - EMIN) * NB / (EMAX - EMIN)
    b = np.trunc(x).astype(int)
    lo = EMIN + ((EMAX - EMIN) * b) / NB
    hi = EMIN + ((EMAX - EMIN) * (b + 1)) / NB
    sb = np.trunc((e - lo) * 10.0 / (hi - lo)).astype(int)
    expect = np.zeros((NB, 10))
    np.add.at(expect, (b, sb), 1.0)
    assert np.all(sub[expect == 0] == 0)
    assert sub.sum() == wl.hist.sum() == n - 0  # every step visits, in range
    tv = 0.5 * np.abs(sub / sub.sum() - expect / expect.sum()).sum()
    assert tv < 0.03, tv


def test_inverse_time_rule_and_convergence(cell):
    upd = _updater(cell, 13)
    wl = _walker(cell, inverse_time=True, f_min=1e-3)
    flats = []
    t = 0
    for _ in range(3000):
        before = (wl.state.n_flat, wl.state.f, wl.state.inv_active, wl.state.since_flat)
        used = []
        wl_run(upd, wl, 7, 1, 1, 1000, move_begin=t, counters=used)
        t += used[0]
        if wl.state.n_flat > before[0]:
            flats.append(dict(f_before=before[1], f=wl.state.f, took=before[3] + 1000, steps=wl.state.steps,
                              active=wl.state.inv_active))
        if wl.state.converged:
            break
    assert wl.state.converged == 1
    act = next(i for i, fl in enumerate(flats) if fl["active"])
    assert act > 0
    for i, fl in enumerate(flats[:act]):
        assert fl["f"] == fl["f_before"] / 2.0
        assert i == 0 or fl["took"] <= flats[i - 1]["took"]
    assert flats[act]["took"] > flats[act - 1]["took"]
    ft = flats[act]["f_before"] * flats[act]["steps"] / NB
    for fl in flats[act:]:
        assert abs(fl["f"] * fl["steps"] / NB - ft) <= 1e-12 * ft
    assert _ln_g_error(cell, wl) < 0.6


def _walk_setup(c, lib_wl):
    """A walker on a high-energy state, every occupied bin known with flat
    counts except a low one, one step per check: the first check cuts the
    window below the walker with at most one recorded state, so the walker
    must walk down into the window."""
    known = (c["g"] > 0).astype(np.uint8)
    kb = np.flatnonzero(known)
    low = int(kb[len(kb) // 2])
    hist = np.where(known, 100000, 0).astype(np.int32)
    hist[low] = 10
    k = int(np.argmax(c["energies"]))
    b = int((c["energies"][k] - EMIN) * NB / (EMAX - EMIN))
    assert b > low + 8
    return k, dict(cur_bin=b, is_first=False, hist=hist, known=known), low


def test_walk_into_a_window_below_the_walker(cell):
    k, preset, low = _walk_setup(cell, None)
    upd = _updater(cell, k)
    wl = _walker(cell, f=1e-3, f_min=-1.0, adaptive=True, min_width=3, **preset)
    used = []
    wl_run(upd, wl, 8, 3, 1, 1, counters=used)
    st = wl.state
    assert st.n_edges == 1 and st.upper == low + 8
    assert wl.saved.sum() <= 1 and st.n_walk > 0
    assert used[0] == 1 + st.n_walk
    assert 0 <= st.cur_bin < st.upper
    e = _energy_of(cell, upd.get_occupation())
    assert int((e - EMIN) * NB / (EMAX - EMIN)) == st.cur_bin
    # the walk went downhill only: every move it kept lowered the energy
    assert e < cell["energies"][k]


# ------------------------------------------------------------------ GPU
def _gpu_batch(c, occ, seed, uid):
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    return ChainBatch([CEModel(c["s"], c["eci"])], [dict(kind=CANONICAL, temperature=1000.0, occupation=occ,
                                                          seed=seed, uid=uid)])


def _compare(batch, chain, wl, ora, log, n_counters, s):
    got = batch.wl_get(chain)
    sch = batch.wl_scheme(chain)
    st = wl.state
    bits = np.unpackbits(log[chain], bitorder="little")
    ref = np.unpackbits(np.asarray(wl.ref_bits, dtype=np.uint8), bitorder="little")
    # both logs keep the bits of the counters that fit (moves back into a
    # window extend the counters past the steps)
    n = min(int(n_counters), bits.size, ref.size)
    bits, ref = bits[:n], ref[:n]
    mism = np.nonzero(bits != ref)[0]
    assert mism.size == 0, "first mismatching counter %d" % mism[0]
    assert got["n_moves"] == st.n_moves and sch["n_walk"] == st.n_walk
    assert np.array_equal(got["hist"], wl.hist)
    assert np.array_equal(got["known"], wl.known)
    assert np.array_equal(got["logdos"], wl.logdos)
    assert got["f"] == st.f and got["current_bin"] == st.cur_bin and got["converged"] == st.converged
    assert (got["n_accept"], got["n_outside"], got["n_self"], got["n_flat"]) == \
        (st.n_accept, st.n_outside, st.n_self, st.n_flat)
    for key in ("inv_active", "n_overlap", "n_edges", "has_first", "steps", "since_flat", "prev_conv"):
        assert sch[key] == getattr(st, key), key
    assert sch["inv_factor"] == st.inv_factor
    assert abs(sch["bin_change"] - st.bin_change) <= 1e-9 * max(1.0, st.bin_change)
    if st.adaptive:
        assert sch["upper"] == st.upper
        assert np.array_equal(sch["sub_bins"].ravel(), wl.sub_hist)
        assert np.array_equal(sch["saved"], wl.saved)
        ne = st.n_edges
        assert np.array_equal(sch["edges"][:ne], wl.edges[:ne])
        assert np.array_equal(sch["edge_n"][:ne], wl.edge_n[:ne])
        vals = wl.edge_vals.reshape(NB, NB)
        for e in range(ne):  # row e holds edge_n[e] values
            assert np.array_equal(sch["edge_vals"][e, :wl.edge_n[e]], vals[e, :wl.edge_n[e]])
    assert np.array_equal(batch.occupation(chain), ora.get_occupation())
    assert abs(batch.energies()[chain] - ora.get_energy()) / s.n_sites <= 1e-10


def _run_pair(c, k, seed, uid, wl_kw, n_checks, check_every, batch=None, chain=0):
    """The oracle walker and the GPU walker `chain` of `batch` (a new batch
    on state k when None) from the same start; returns (batch, oracle
    walker, oracle updater, counters used)."""
    ora = _updater(c, k)
    wl = _walker(c, **wl_kw)
    used = []
    wl.ref_bits = wl_run(ora, wl, seed, uid, n_checks, check_every, counters=used)
    if batch is None:
        batch = _gpu_batch(c, c["states"][k], seed, uid)
    kw = dict(wl_kw)
    scheme = {key: kw.pop(key) for key in ("inverse_time", "adaptive", "min_width") if key in kw}
    for key in ("f", "f_min", "flatness", "cur_bin"):
        kw.pop(key, None)
    init = wl_kw.get("cur_bin", 0)
    batch.wl_set(chain, NB, EMIN, EMAX, wl_kw.get("f", 1.0), f_min=wl.state.f_min, flatness=wl.state.flatness,
                 current_bin=init, **kw)
    batch.wl_set_scheme(chain, **scheme)
    return batch, wl, ora, used[0]


def _sub_edges_clear(c):
    """No energy level within 1e-6 of a sub-bin edge (the running energy on
    the GPU and the oracle's recomputed one differ in the last bits)."""
    e = c["energies"]
    x = (e - EMIN) * NB / (EMAX - EMIN)
    b = np.trunc(x).astype(int)
    lo = EMIN + ((EMAX - EMIN) * b) / NB
    hi = EMIN + ((EMAX - EMIN) * (b + 1)) / NB
    y = (e - lo) * 10.0 / (hi - lo)
    return np.min(np.abs(y - np.round(y)))


@pytest.mark.gpu
@pytest.mark.parametrize("n_checks", [40, 3000])
def test_gpu_adaptive_walker_matches_oracle(gpu, cell, n_checks):
    """Mid-run (windows cut, not yet joined) and to convergence."""
    assert _sub_edges_clear(cell) > 1e-6
    print("adaptive walker", n_checks, flush=True)
    batch, wl, ora, n = _run_pair(cell, 101, 1, 3, dict(adaptive=True, min_width=3), n_checks, 1000)
    log, _ = batch.wl_run(n_checks, 1000, log=True)
    _compare(batch, 0, wl, ora, log, n, cell["s"])
    if n_checks == 3000:
        assert wl.state.converged == 1 and wl.state.has_first == 1
    else:
        assert wl.state.n_edges > 0


@pytest.mark.gpu
@pytest.mark.parametrize("adaptive", [False, True])
def test_gpu_inverse_time_walker_matches_oracle(gpu, cell, adaptive):
    print("inverse-time walker, adaptive", adaptive, flush=True)
    batch, wl, ora, n = _run_pair(cell, 13, 7, 1, dict(inverse_time=True, adaptive=adaptive, min_width=3,
                                                       f_min=2e-3), 600, 1000)
    log, _ = batch.wl_run(600, 1000, log=True)
    _compare(batch, 0, wl, ora, log, n, cell["s"])
    assert wl.state.inv_active == 1


@pytest.mark.gpu
def test_gpu_walk_into_window_matches_oracle(gpu, cell):
    k, preset, low = _walk_setup(cell, None)
    print("walk into window", flush=True)
    batch, wl, ora, n = _run_pair(cell, k, 8, 3, dict(f=1e-3, f_min=-1.0, adaptive=True, min_width=3, **preset),
                                  1, 1)
    log, _ = batch.wl_run(1, 1, log=True)
    assert wl.state.n_walk > 0
    _compare(batch, 0, wl, ora, log, n, cell["s"])


@pytest.mark.gpu
def test_gpu_mixed_schemes_in_one_batch(gpu, cell):
    """Fixed f, inverse time and adaptive windows on three walkers of one
    launch, continued over a second launch."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    model = CEModel(cell["s"], cell["eci"])
    ks = (21, 22, 23)
    batch = ChainBatch([model] * 3, [dict(kind=CANONICAL, temperature=1000.0, occupation=cell["states"][k],
                                          seed=5, uid=10 + i) for i, k in enumerate(ks)])
    kws = (dict(), dict(inverse_time=True), dict(adaptive=True, min_width=3))
    pairs = []
    for i, (k, kw) in enumerate(zip(ks, kws)):
        _, wl, ora, _ = _run_pair(cell, k, 5, 10 + i, dict(kw, f_min=1e-3), 0, 1000, batch=batch, chain=i)
        pairs.append((wl, ora))
    print("mixed schemes", flush=True)
    for rep in range(2):
        log, _ = batch.wl_run(150, 1000, log=True)
        for i, (wl, ora) in enumerate(pairs):
            used = []
            move_begin = int(wl.state.n_moves + wl.state.n_walk)
            wl.ref_bits = wl_run(ora, wl, 5, 10 + i, 150, 1000, move_begin=move_begin, counters=used)
            _compare(batch, i, wl, ora, log, used[0], cell["s"])


@pytest.mark.gpu
def test_python_driver_adaptive_windows(gpu, cell, tmp_path):
    """WangLandau with the CE calculator in adaptive-window mode: eight
    walkers to f < 1e-5 through several cuts; their averaged ln g matches
    the enumerated one, and the sub-bin counts are written as CSV."""
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.wang_landau import WangLandau
    s = cell["s"]
    occ = cell["states"][100]
    atoms = SimpleAtoms([s.species[v] for v in occ])
    CE(atoms, s, eci=dict(cell["eci"]))
    wl = WangLandau(atoms, Nbins=NB, initial_f=1.0, flatness_criteria=0.8, fmin=1e-5, Emin=EMIN, Emax=EMAX,
                    check_convergence_every=1000, n_walkers=8, seed=6)
    out = tmp_path / "subbin.csv"
    print("python driver, adaptive windows", flush=True)
    assert wl.run_fast_sampler(maxsteps=3000000, mode="adaptive_windows", minimum_window_width=3,
                               sub_bin_file=str(out))
    assert all(wl.scheme_state(w)["has_first"] == 1 for w in range(8))
    g = cell["g"]
    known = g > 0
    est = wl.logdos()
    assert np.array_equal(~np.isnan(est), known)
    d = est[known] - np.log(g[known])
    d -= d.mean()
    # eight walkers' windows join at different bins; their mean is closer
    # than one walker's bound
    assert np.abs(d).max() < 1.0, d
    rows = out.read_text().splitlines()
    assert len(rows) == NB and all(r.count(",") == 10 for r in rows)
    assert sum(int(v) for r in rows for v in r.split(",") if v) == int(wl.sub_bin_distribution(0).sum())
    wl.close()


# ------------------------------------------------------- shared histogram
@pytest.mark.gpu
@pytest.mark.parametrize("inverse_time", [False, True])
def test_gpu_shared_histogram_one_walker_matches_oracle(gpu, cell, inverse_time):
    """One walker on a shared histogram makes the same moves, visits and f
    updates as the oracle's own-histogram walker (nothing races)."""
    print("shared histogram, one walker, inverse time", inverse_time, flush=True)
    batch, wl, ora, n = _run_pair(cell, 31, 2, 4, dict(inverse_time=inverse_time, f_min=1e-3), 400, 1000)
    log, _ = batch.wl_run(400, 1000, log=True, shared=True)
    got = batch.wl_get(0)
    st = wl.state
    bits = np.unpackbits(log[0], bitorder="little")[:n]
    ref = np.unpackbits(np.asarray(wl.ref_bits, dtype=np.uint8), bitorder="little")[:n]
    assert np.array_equal(bits, ref)
    assert np.array_equal(got["hist"], wl.hist) and np.array_equal(got["known"], wl.known)
    assert np.array_equal(got["logdos"], wl.logdos)
    assert got["f"] == st.f and got["converged"] == st.converged and got["n_flat"] == st.n_flat
    assert got["n_moves"] == st.n_moves and got["current_bin"] == st.cur_bin
    sch = batch.wl_scheme(0)
    assert (sch["inv_active"], sch["steps"], sch["prev_conv"]) == (st.inv_active, st.steps, st.prev_conv)
    assert np.array_equal(batch.occupation(0), ora.get_occupation())


@pytest.mark.gpu
def test_gpu_shared_histogram_with_fixed_ln_g_adds_the_walkers_visits(gpu, cell):
    """f = 0 and never flat: ln g stays as given, so each walker's moves are
    its own oracle walker's and the shared counts are the sum of theirs."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    g = cell["g"]
    lg = np.log(np.where(g > 0, g, 1.0))
    ks = list(range(40, 48))
    W = len(ks)
    batch = ChainBatch([CEModel(cell["s"], cell["eci"])] * W,
                       [dict(kind=CANONICAL, temperature=1000.0, occupation=cell["states"][k], seed=9, uid=k)
                        for k in ks])
    for w in range(W):
        batch.wl_set(w, NB, EMIN, EMAX, 0.0, f_min=-1.0, flatness=2.0, logdos=lg if w == 0 else None)
    per, n_checks = 250, 40
    print("shared histogram, fixed ln g,", W, "walkers", flush=True)
    log, _ = batch.wl_run(n_checks, per * W, log=True, shared=True)
    hist = np.zeros(NB, dtype=np.int64)
    for w, k in enumerate(ks):
        ora = _updater(cell, k)
        wl = OracleWL(NB, EMIN, EMAX, 0.0, f_min=-1.0, flatness=2.0, logdos=lg)
        ref = wl_run(ora, wl, 9, k, n_checks, per)
        n = n_checks * per
        assert np.array_equal(np.unpackbits(log[w], bitorder="little")[:n],
                              np.unpackbits(ref, bitorder="little")[:n]), w
        assert np.array_equal(batch.occupation(w), ora.get_occupation())
        assert batch.wl_get(w)["n_accept"] == wl.state.n_accept
        hist += wl.hist
    got = batch.wl_get(0)
    assert np.array_equal(got["hist"], hist)
    assert np.array_equal(got["logdos"], lg)


@pytest.mark.gpu
def test_gpu_shared_histogram_converges_to_exact_dos(gpu, cell):
    """Sixteen walkers on one histogram, f from 1 to below 1e-5."""
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.wang_landau import WangLandau
    s = cell["s"]
    atoms = SimpleAtoms([s.species[v] for v in cell["states"][100]])
    CE(atoms, s, eci=dict(cell["eci"]))
    wl = WangLandau(atoms, Nbins=NB, initial_f=1.0, flatness_criteria=0.8, fmin=1e-5, Emin=EMIN, Emax=EMAX,
                    check_convergence_every=16000, n_walkers=16, seed=12, shared_histogram=True)
    print("shared histogram, 16 walkers", flush=True)
    assert wl.run_fast_sampler(maxsteps=16000 * 400)
    assert wl.f < 1e-5
    g = cell["g"]
    known = g > 0
    est = wl.logdos()
    assert np.array_equal(~np.isnan(est), known)
    d = est[known] - np.log(g[known])
    d -= d.mean()
    assert np.abs(d).max() < 0.3, d
    wl.close()


@pytest.mark.gpu
@pytest.mark.parametrize("crystal", ["hcp_ternary", "sg217_binary", "fcc_quaternary_tables_in_hbm"])
def test_gpu_scheme_walkers_on_crystals(gpu, crystal):
    """Adaptive windows with the inverse-time law on crystals with sites
    not related by translations: HCP (two sites per cell, one translation
    group, ternary) and space group I-43m (29 translation groups, records
    read from LDS); and on quaternary FCC 8^3 with clusters up to 4-body to
    the 3rd shell, whose energy tables (three values per code side by side)
    are too large for LDS; walkers equal the oracle bit for bit."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.crystal import HCPSettings, Spacegroup217Settings
    from cemc_amd.lattice import random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    n_checks, per = 40, 800
    if crystal == "hcp_ternary":
        s = HCPSettings((4, 4, 3), ["Al", "Mg", "Si"], {2: 1.45, 3: 1.05})
        n_sp = 3
    elif crystal == "fcc_quaternary_tables_in_hbm":
        s = FCCSettings(8, ["A", "B", "C", "D"], {2: 3, 3: 3, 4: 3})
        n_sp = 4
        n_checks, per = 8, 400  # 855 correlation functions: the oracle's updates are slow
    else:
        s = Spacegroup217Settings(1, 10.553, [(0, 0, 0), (0.324, 0.324, 0.324), (0.3582, 0.3582, 0.0393),
                                              (0.0954, 0.0954, 0.2725)], [["Al", "Mg"]] * 4, {2: 4.1, 3: 4.1})
        n_sp = 2
    names = s.cf_names()
    eci = seeded_ecis(s, 21, 0.01)
    occ = random_occupation(s.n_sites, n_sp, 14)
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    e0 = ora.get_energy()
    nb, half = 40, 0.6
    w = 2.0 * half / nb
    lo = e0 - half + w / 3.0
    wl = OracleWL(nb, lo, lo + 2.0 * half, 1.0, f_min=1e-3, flatness=0.6, cur_bin=20, inverse_time=True,
                  adaptive=True, min_width=3, n_sites=s.n_sites, n_cf=len(names))
    used = []
    ref = wl_run(ora, wl, 3, 2, n_checks, per, counters=used)
    batch = ChainBatch([CEModel(s, eci)], [dict(kind=CANONICAL, temperature=1000.0, occupation=occ, seed=3, uid=2)])
    batch.wl_set(0, nb, lo, lo + 2.0 * half, 1.0, f_min=1e-3, flatness=0.6, current_bin=20)
    batch.wl_set_scheme(0, inverse_time=True, adaptive=True, min_width=3)
    print("scheme walker on", crystal, flush=True)
    log, _ = batch.wl_run(n_checks, per, log=True)
    got = batch.wl_get(0)
    sch = batch.wl_scheme(0)
    st = wl.state
    n = min(used[0], n_checks * per)
    assert np.array_equal(np.unpackbits(log[0], bitorder="little")[:n], np.unpackbits(ref, bitorder="little")[:n])
    assert np.array_equal(got["hist"], wl.hist) and np.array_equal(got["logdos"], wl.logdos)
    assert got["f"] == st.f and got["current_bin"] == st.cur_bin and got["n_flat"] == st.n_flat
    assert (sch["n_edges"], sch["upper"], sch["n_walk"], sch["inv_active"]) == (st.n_edges, st.upper, st.n_walk,
                                                                                  st.inv_active)
    assert np.array_equal(sch["sub_bins"].ravel(), wl.sub_hist)
    assert np.array_equal(batch.occupation(0), ora.get_occupation())


# -------------------------------------------- semi-grand-canonical walkers
NB_S, EMIN_S = 38, -0.09275  # 2x2x3 cell, all 4096 states: no level within 1.7 % of a bin edge
EMAX_S = EMIN_S + NB_S * 0.005


@pytest.fixture(scope="module")
def sgc_cell():
    """2x2x3 binary cell with every composition (4096 states); the singlet
    ECI carries a chemical potential of 0.008 eV."""
    s = FCCSettings((2, 2, 3), ["A", "B"], {2: 2, 3: 1})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    c2 = [n for n in names if n.startswith("c2")]
    c3 = [n for n in names if n.startswith("c3")]
    eci[c2[0]], eci[c2[1]], eci[c3[0]] = 0.01, -0.004, 0.003
    eci["c1_0"] = 0.003 - 0.008
    N = s.n_sites
    states = [np.array(c, dtype=np.int8) for c in itertools.product([0, 1], repeat=N)]
    energies = np.array([N * sum(eci[n] * v for n, v in brute_force_cf(s, st, names).items()) for st in states])
    x = (energies - EMIN_S) * NB_S / (EMAX_S - EMIN_S)
    assert np.min(np.abs(x - np.round(x))) > 0.015
    g = np.bincount(np.trunc(x).astype(int), minlength=NB_S).astype(float)
    return dict(s=s, names=names, eci=eci, states=states, energies=energies, g=g, model=OracleModel(s, names))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sgc_walker_converges_to_exact_dos(sgc_cell, seed):
    """Flips over all compositions: ln g of the energy with the chemical
    potential converges to the enumerated one; every occupied bin becomes
    known."""
    c = sgc_cell
    occ = c["states"][100 + 37 * seed]
    upd = OracleUpdater(c["model"], occ, brute_force_cf(c["s"], occ, c["names"]), c["eci"])
    wl = OracleWL(NB_S, EMIN_S, EMAX_S, 1.0, f_min=1e-4, flatness=0.8, kind=1)
    wl_run(upd, wl, seed, 1, 400, 2000)
    g = c["g"]
    known = wl.known != 0
    assert wl.state.converged == 1 and np.array_equal(known, g > 0)
    d = (wl.logdos - np.log(np.where(g > 0, g, 1.0)))[known]
    # ln g spans ln(528) = 6.3
    assert np.abs(d - d.mean()).max() < 0.35


def test_sgc_walker_with_fixed_ln_g_samples_states_uniformly(sgc_cell):
    """f = 0 and ln g = 0: flips accept everything in range, so visits
    follow g (every state equally likely), and the composition changes."""
    c = sgc_cell
    occ = c["states"][77]
    upd = OracleUpdater(c["model"], occ, brute_force_cf(c["s"], occ, c["names"]), c["eci"])
    wl = OracleWL(NB_S, EMIN_S, EMAX_S, 0.0, f_min=-1.0, flatness=2.0, kind=1)
    n = 100 * 2000
    wl_run(upd, wl, 5, 2, 100, 2000)
    p = c["g"] / c["g"].sum()
    assert wl.hist.sum() == n and wl.state.n_outside == 0
    assert 0.5 * np.abs(wl.hist / n - p).sum() < 0.03
    assert upd.get_occupation().sum() != occ.sum() or wl.state.n_accept > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["binary_cell", "ternary_periodic", "ternary_adaptive"])
def test_gpu_sgc_walkers_match_oracle(gpu, sgc_cell, case):
    from cemc_amd.batch import SGC, ChainBatch
    from cemc_amd.lattice import random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    if case == "binary_cell":
        s, eci, occ = sgc_cell["s"], dict(sgc_cell["eci"]), sgc_cell["states"][100]
        mu = {"c1_0": 0.0}
        nb, lo, hi = NB_S, EMIN_S, EMAX_S
        kw = dict(f_min=1e-3)
    else:
        s = FCCSettings(5, ["A", "B", "C"], {2: 3, 3: 2})
        eci = seeded_ecis(s, 8, 0.01)
        occ = random_occupation(s.n_sites, 3, 6)
        mu = {"c1_0": 0.02, "c1_1": -0.01}
        names = s.cf_names()
        e_mu = dict(eci)
        for k, v in mu.items():
            e_mu[k] -= v
        e0 = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), e_mu).get_energy()
        nb, half = 50, 0.5
        w = 2.0 * half / nb
        lo = e0 - half + w / 3.0
        hi = lo + 2.0 * half
        kw = dict(f_min=1e-3, flatness=0.6, adaptive=case == "ternary_adaptive", min_width=3, inverse_time=True)
    names = s.cf_names()
    e_mu = dict(eci)
    for k, v in mu.items():
        e_mu[k] -= v
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), e_mu)
    wl = OracleWL(nb, lo, hi, 1.0, kind=1, n_sites=s.n_sites, n_cf=len(names), cur_bin=nb // 2,
                  **{k: v for k, v in kw.items() if k != "f_min"}, f_min=kw["f_min"])
    used = []
    ref = wl_run(ora, wl, 4, 9, 30, 1000, counters=used)
    batch = ChainBatch([CEModel(s, eci)], [dict(kind=SGC, temperature=1000.0, occupation=occ, seed=4, uid=9,
                                                chem_pot=mu)])
    assert abs(batch.energies()[0] - OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names),
                                                   e_mu).get_energy()) < 1e-9
    batch.wl_set(0, nb, lo, hi, 1.0, f_min=kw["f_min"], flatness=wl.state.flatness, current_bin=nb // 2)
    if case != "binary_cell":
        batch.wl_set_scheme(0, inverse_time=True, adaptive=kw["adaptive"], min_width=3)
    print("SGC walker", case, flush=True)
    log, _ = batch.wl_run(30, 1000, log=True)
    got = batch.wl_get(0)
    st = wl.state
    n = min(used[0], 30 * 1000)
    assert np.array_equal(np.unpackbits(log[0], bitorder="little")[:n], np.unpackbits(ref, bitorder="little")[:n])
    assert np.array_equal(got["hist"], wl.hist) and np.array_equal(got["logdos"], wl.logdos)
    assert got["f"] == st.f and got["current_bin"] == st.cur_bin and got["n_accept"] == st.n_accept
    assert np.array_equal(batch.occupation(0), ora.get_occupation())
    assert abs(batch.energies()[0] - ora.get_energy()) <= 1e-10 * s.n_sites
    if case == "ternary_adaptive":
        sch = batch.wl_scheme(0)
        assert np.array_equal(sch["sub_bins"].ravel(), wl.sub_hist) and sch["upper"] == st.upper


@pytest.mark.gpu
def test_gpu_sgc_walkers_on_a_shared_histogram_add_their_visits(gpu, sgc_cell):
    """Semi-grand-canonical walkers on one histogram with f = 0 and flatness
    never reached: each walker flips as its own oracle walker does, and the
    shared counts are the sum of theirs."""
    from cemc_amd.batch import SGC, ChainBatch
    from cemc_amd.model import CEModel
    c = sgc_cell
    g = c["g"]
    lg = np.log(np.where(g > 0, g, 1.0))
    ks = [100, 700, 1500, 2900]
    W = len(ks)
    batch = ChainBatch([CEModel(c["s"], dict(c["eci"]))] * W,
                       [dict(kind=SGC, temperature=1000.0, occupation=c["states"][k], seed=6, uid=k,
                             chem_pot={"c1_0": 0.0}) for k in ks])
    for w in range(W):
        batch.wl_set(w, NB_S, EMIN_S, EMAX_S, 0.0, f_min=-1.0, flatness=2.0, logdos=lg if w == 0 else None)
    per, n_checks = 300, 20
    print("shared histogram, SGC,", W, "walkers", flush=True)
    log, _ = batch.wl_run(n_checks, per * W, log=True, shared=True)
    hist = np.zeros(NB_S, dtype=np.int64)
    for w, k in enumerate(ks):
        occ = c["states"][k]
        ora = OracleUpdater(c["model"], occ, brute_force_cf(c["s"], occ, c["names"]), c["eci"])
        wl = OracleWL(NB_S, EMIN_S, EMAX_S, 0.0, f_min=-1.0, flatness=2.0, logdos=lg, kind=1)
        ref = wl_run(ora, wl, 6, k, n_checks, per)
        n = n_checks * per
        assert np.array_equal(np.unpackbits(log[w], bitorder="little")[:n],
                              np.unpackbits(ref, bitorder="little")[:n]), w
        assert np.array_equal(batch.occupation(w), ora.get_occupation())
        hist += wl.hist
    got = batch.wl_get(0)
    assert np.array_equal(got["hist"], hist)
    assert np.array_equal(got["logdos"], lg)


@pytest.mark.gpu
def test_python_driver_sgc_converges_to_exact_dos(gpu, sgc_cell):
    """WangLandau in the semi-grand-canonical ensemble with the CE
    calculator: eight walkers; the averaged ln g matches the enumerated one
    of the energy with the chemical potential."""
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.wang_landau import WangLandau
    c = sgc_cell
    s = c["s"]
    eci = dict(c["eci"])
    eci["c1_0"] += 0.008  # the calculator's own ECI; the walkers fold mu = 0.008 back in
    atoms = SimpleAtoms([s.species[v] for v in c["states"][300]])
    CE(atoms, s, eci=eci)
    wl = WangLandau(atoms, Nbins=NB_S, initial_f=1.0, flatness_criteria=0.8, fmin=1e-4, Emin=EMIN_S, Emax=EMAX_S,
                    check_convergence_every=2000, n_walkers=8, seed=3, ensemble="semi-grand-canonical",
                    chem_pot={"c1_0": 0.008})
    print("python driver, SGC", flush=True)
    assert wl.run_fast_sampler(maxsteps=800000)
    g = c["g"]
    known = g > 0
    est = wl.logdos()
    assert np.array_equal(~np.isnan(est), known)
    d = est[known] - np.log(g[known])
    d -= d.mean()
    assert np.abs(d).max() < 0.25, d
    wl.close()
