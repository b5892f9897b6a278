"""Wang-Landau walkers (cpp/src/wang_landau_sampler.cpp:232-404,
cpp/src/histogram.cpp:66-172).

CPU (the oracle's restatement, pinned by what the mathematics and the
reference's rules fix):
* on a 2x2x3 cell whose 924 states at fixed composition are enumerated, the
  walker's ln g converges to the exact ln g per bin, and it visits exactly
  the bins that hold states;
* with f = 0 the walk is a Markov chain with acceptance
  min(1, g~(old) / g~(new)) for the fixed ln g~: for ln g~ = 0 the visit counts
  follow the exact g, for ln g~ = exact ln g they are flat;
* every in-range step counts one visit; proposals above the range count a
  visit of the current bin, proposals below it do not, and nothing is
  counted before the first in-range proposal (wang_landau_sampler.cpp:251-260);
* the bin is the truncation toward zero of (E - Emin) Nbins / (Emax - Emin),
  so energies up to one bin below Emin fall in bin 0 (histogram.cpp:66-69);
* flatness looks at known bins only, compares strictly, halves f and
  clears the counts; a walker stops at the first check with f < fmin.

GPU: the kernel reproduces the oracle walker exactly (accept bits, visit
counts, known bins, ln g bitwise, f, current bin, counters, lattice) on
table and closed-form lattices, two to four species, records in registers
and in LDS, several walkers of different models in one batch, and the
range and flatness cases above.
"""
import itertools

import numpy as np
import pytest

from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
from oracle.brute_force import brute_force_cf
from oracle.oracle import OracleModel, OracleUpdater, OracleWL, wl_run

E_TOL = 1e-10  # eV per site


def _bits(packed, n):
    return np.unpackbits(np.asarray(packed, dtype=np.uint8), bitorder="little")[:n]


# ------------------------------------------------------------- tiny cell
NB, EMIN, EMAX = 14, -0.0905, 0.0495  # bin width 0.01; no level within 1/60 of a bin edge


@pytest.fixture(scope="module")
def tiny():
    """2x2x3 binary cell, 6 + 6 atoms: every state and its energy."""
    s = FCCSettings((2, 2, 3), ["A", "B"], {2: 2, 3: 1})
    names = s.cf_names()
    eci = {n: 0.0 for n in names}
    c2 = [n for n in names if n.startswith("c2")]
    c3 = [n for n in names if n.startswith("c3")]
    eci[c2[0]], eci[c2[1]], eci[c3[0]] = 0.01, -0.004, 0.003
    N = s.n_sites
    states = [np.array(c, dtype=np.int8) for c in itertools.product([0, 1], repeat=N) if sum(c) == N // 2]
    energies = np.array([N * sum(eci[n] * v for n, v in brute_force_cf(s, st, names).items() if n in eci)
                         for st in states])
    x = (energies - EMIN) * NB / (EMAX - EMIN)
    assert np.min(np.abs(x - np.round(x))) > 0.01
    g = np.bincount(np.trunc(x).astype(int), minlength=NB).astype(float)
    return dict(s=s, names=names, eci=eci, states=states, energies=energies, g=g, model=OracleModel(s, names))


def _updater(t, k):
    occ = t["states"][k]
    return OracleUpdater(t["model"], occ, brute_force_cf(t["s"], occ, t["names"]), t["eci"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_converges_to_exact_dos(tiny, seed):
    upd = _updater(tiny, 5 + seed)
    wl = OracleWL(NB, EMIN, EMAX, 1.0, f_min=1e-4, flatness=0.8)
    wl_run(upd, wl, seed, 1, 200, 2000)
    st = wl.state
    assert st.converged == 1 and st.n_flat == 14 and st.f == 2.0 ** -14
    assert st.n_moves % 2000 == 0 and st.n_moves < 200 * 2000
    g = tiny["g"]
    known = wl.known != 0
    assert np.array_equal(known, g > 0)
    d = (wl.logdos - np.log(np.where(g > 0, g, 1.0)))[known]
    d -= d.mean()
    # exact ln g spans 4.1; the fixed-f estimate at f = 1e-4 is within 0.3
    assert np.abs(d).max() < 0.3, d


def test_oracle_fixed_ln_g_walks_follow_exact_weights(tiny):
    """f = 0: ln g~ = 0 samples states uniformly (visits follow g); ln g~ =
    ln g samples bins uniformly (flat visits)."""
    g = tiny["g"]
    occ_bins = g > 0
    p = g / g.sum()
    n = 100 * 2000
    upd = _updater(tiny, 3)
    wl = OracleWL(NB, EMIN, EMAX, 0.0, f_min=-1.0, flatness=2.0)
    wl_run(upd, wl, 4, 2, 100, 2000)
    assert wl.state.n_moves == n and wl.state.n_outside == 0
    assert wl.hist.sum() == n and np.all(wl.logdos == 0.0)
    assert np.array_equal(wl.known != 0, occ_bins)
    assert 0.5 * np.abs(wl.hist / n - p).sum() < 0.03
    upd = _updater(tiny, 3)
    lg = np.log(np.where(occ_bins, g, 1.0))
    wl = OracleWL(NB, EMIN, EMAX, 0.0, f_min=-1.0, flatness=2.0, logdos=lg)
    wl_run(upd, wl, 4, 3, 100, 2000)
    share = wl.hist[occ_bins] / n
    assert np.abs(share * occ_bins.sum() - 1.0).max() < 0.15, share
    # the detailed-balance weights really differ between the two: the
    # uniform-state walk is far from flat
    assert np.abs(p[occ_bins] * occ_bins.sum() - 1.0).max() > 1.0


def test_oracle_visit_counting_in_and_out_of_range(tiny):
    e = tiny["energies"]
    n = 5 * 1000
    # every state in range: one visit per move, ln g grows by f per visit
    wl = OracleWL(NB, EMIN, EMAX, 1e-3, f_min=0.0, flatness=2.0)
    wl_run(_updater(tiny, 7), wl, 1, 1, 5, 1000)
    assert wl.hist.sum() == n and abs(wl.logdos.sum() - 1e-3 * n) < 1e-9
    # every state below the range by more than a bin: nothing is counted
    w = (EMAX - EMIN) / NB
    lo = e.max() + 2 * w
    wl = OracleWL(NB, lo, lo + NB * w, 1.0, f_min=0.0, flatness=2.0, cur_bin=3, is_first=False)
    up = _updater(tiny, 7)
    occ0 = up.get_occupation().copy()
    wl_run(up, wl, 1, 1, 5, 1000)
    assert wl.hist.sum() == 0 and wl.state.n_outside == n and wl.state.n_accept == 0
    assert np.array_equal(up.get_occupation(), occ0)
    # every state above the range: each move counts the current bin once
    hi = e.min() - 2 * w
    wl = OracleWL(NB, hi - NB * w, hi, 1.0, f_min=0.0, flatness=2.0, cur_bin=3, is_first=False)
    wl_run(_updater(tiny, 7), wl, 1, 1, 5, 1000)
    assert wl.hist[3] == n and wl.hist.sum() == n and wl.logdos[3] == float(n)
    assert wl.state.n_outside == n
    # the same before the first in-range proposal: nothing is counted
    wl = OracleWL(NB, hi - NB * w, hi, 1.0, f_min=0.0, flatness=2.0, cur_bin=3, is_first=True)
    wl_run(_updater(tiny, 7), wl, 1, 1, 5, 1000)
    assert wl.hist.sum() == 0 and not wl.known.any()


def test_oracle_bin_truncates_toward_zero(tiny):
    """Emin half a bin above the ground level: ground states fall in bin 0
    (truncation of -0.5), alone there; with floor they would be outside."""
    e = tiny["energies"]
    w = 0.01
    lo = e.min() + 0.5 * w
    second = np.unique(np.round(e, 9))[1]
    assert second > lo + 3 * w
    wl = OracleWL(30, lo, lo + 30 * w, 0.0, f_min=-1.0, flatness=2.0)
    wl_run(_updater(tiny, 2), wl, 9, 1, 100, 2000)
    assert wl.known[0] == 1 and wl.hist[0] > 0
    assert not wl.known[1:4].any()


def test_oracle_flatness_rule_and_stop(tiny):
    upd = _updater(tiny, 4)
    occ_bins = tiny["g"] > 0
    # preset flat counts on the known bins, one unknown bin at zero: flat
    # (unknown bins are ignored); f halves, counts clear, and f < fmin stops
    # the walker after the first check
    hist = np.where(occ_bins, 100000, 0).astype(np.int32)
    wl = OracleWL(NB, EMIN, EMAX, 1.0, f_min=0.6, flatness=0.8, hist=hist, known=occ_bins.astype(np.uint8))
    wl_run(upd, wl, 3, 1, 10, 500)
    assert wl.state.n_flat == 1 and wl.state.f == 0.5 and wl.state.converged == 1
    assert wl.state.n_moves == 500 and wl.hist.sum() == 0
    # equal counts (every proposal below the range, so none is added) are
    # not flat at criteria 1 (strict comparison) and flat just below it
    lo = tiny["energies"].max() + 0.02
    for crit, flat in ((1.0, 0), (0.9999, 1)):
        wl = OracleWL(NB, lo, lo + 0.14, 0.0, f_min=-1.0, flatness=crit, cur_bin=3, is_first=False, hist=hist,
                      known=occ_bins.astype(np.uint8))
        wl_run(_updater(tiny, 4), wl, 3, 1, 1, 10)
        assert wl.state.n_flat == flat and wl.state.n_outside == 10
    # one known bin far below the mean: not flat
    h2 = hist.copy()
    h2[np.argmax(occ_bins)] = 1000
    wl = OracleWL(NB, EMIN, EMAX, 1.0, f_min=0.0, flatness=0.8, hist=h2, known=occ_bins.astype(np.uint8))
    wl_run(_updater(tiny, 4), wl, 3, 1, 1, 100)
    assert wl.state.n_flat == 0 and wl.state.f == 1.0
    # a converged walker makes no moves
    wl.state.converged = 1
    wl_run(_updater(tiny, 4), wl, 3, 1, 2, 100)
    assert wl.state.n_moves == 100


# ------------------------------------------------------------------ GPU
def _centred(e0, half, n_bins):
    """A range of 2 half around e0 whose bin edges miss e0 by a third of a
    bin: states degenerate with the start (its translations) then never sit
    on an edge, where rounding of the running energy could pick either bin."""
    w = 2.0 * half / n_bins
    lo = e0 - half + w / 3.0
    return dict(n_bins=n_bins, e_min=lo, e_max=lo + 2.0 * half)


def _gpu_vs_oracle(s, eci, occ, wl_args, n_checks, check_every, seed=17, uid=5):
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    wl = OracleWL(**wl_args)
    ref_bits = wl_run(ora, wl, seed, uid, n_checks, check_every)
    batch = ChainBatch([CEModel(s, eci)], [dict(kind=CANONICAL, temperature=1000.0, occupation=occ, seed=seed,
                                                uid=uid)])
    a = dict(wl_args)
    batch.wl_set(0, a.pop("n_bins"), a.pop("e_min"), a.pop("e_max"), a.pop("f"), current_bin=a.pop("cur_bin", 0),
                 **a)
    log, _ = batch.wl_run(n_checks, check_every, log=True)
    got = batch.wl_get(0)
    st = wl.state
    n = int(st.n_moves)
    assert got["n_moves"] == n
    mism = np.nonzero(_bits(log[0], n) != _bits(ref_bits, n))[0]
    assert mism.size == 0, "first mismatching move %d" % mism[0]
    assert np.array_equal(got["hist"], wl.hist)
    assert np.array_equal(got["known"], wl.known)
    assert np.array_equal(got["logdos"], wl.logdos)  # the same additions in the same order
    assert got["f"] == st.f and got["current_bin"] == st.cur_bin and got["is_first"] == st.is_first
    assert got["converged"] == st.converged
    assert (got["n_accept"], got["n_outside"], got["n_self"], got["n_flat"]) == \
        (st.n_accept, st.n_outside, st.n_self, st.n_flat)
    assert np.array_equal(batch.occupation(0), ora.get_occupation())
    assert abs(batch.energies()[0] - ora.get_energy()) / s.n_sites <= E_TOL
    return got, batch


@pytest.mark.gpu
def test_gpu_walker_matches_oracle_tiny_cell(gpu, tiny):
    occ = tiny["states"][5]
    print("tiny cell walker", flush=True)
    got, _ = _gpu_vs_oracle(tiny["s"], tiny["eci"], occ,
                            dict(n_bins=NB, e_min=EMIN, e_max=EMAX, f=1.0, f_min=1e-4, flatness=0.8), 200, 2000)
    assert got["converged"] == 1 and got["n_flat"] == 14


@pytest.mark.gpu
@pytest.mark.parametrize("species,shells,size", [
    (["Al", "Mg"], {2: 3, 3: 2}, 6),
    (["Al", "Mg", "Si"], {2: 3, 3: 2}, 5),
    (["Al", "Mg", "Si", "Cu"], {2: 2, 3: 1}, 5),
    (["Al", "Mg"], {2: 2, 3: 1}, (2, 2, 3)),
])
def test_gpu_walker_matches_oracle(gpu, species, shells, size):
    """Periodic closed-form lattices with two, three and four species, and a
    translation-table cell; the range is centred on the start so proposals
    leave it on both sides."""
    s = FCCSettings(size, species, shells)
    eci = seeded_ecis(s, 31, 0.01)
    occ = random_occupation(s.n_sites, len(species), 12)
    names = s.cf_names()
    e0 = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci).get_energy()
    print("walker", species, shells, size, flush=True)
    _gpu_vs_oracle(s, eci, occ, dict(_centred(e0, 0.4, 50), f=1.0, f_min=1e-3, flatness=0.6, cur_bin=25), 12, 1500)


@pytest.mark.gpu
def test_gpu_walker_records_in_lds(gpu):
    """More sub-clusters than eight records per lane: records read from LDS."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    s = FCCSettings(8, ["A", "B", "C"], {2: 5, 3: 4, 4: 3})
    eci = seeded_ecis(s, 5, 0.01)
    occ = random_occupation(s.n_sites, 3, 2)
    m = CEModel(s, eci)
    probe = ChainBatch([m], [dict(kind=CANONICAL, temperature=900.0, occupation=occ, seed=1, uid=0)])
    probe.set_waves(1)
    # one wave cannot hold the records in registers (more than 8 per lane),
    # so the sweep falls back to more waves and the walker reads them from LDS
    assert probe.launch_shape(0)[0] > 1
    probe.close()
    names = s.cf_names()
    e0 = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci).get_energy()
    print("records in LDS walker", flush=True)
    _gpu_vs_oracle(s, eci, occ, dict(_centred(e0, 0.3, 40), f=1.0, f_min=1e-3, flatness=0.6, cur_bin=20), 4, 800)


@pytest.mark.gpu
def test_gpu_walker_range_and_flatness_cases(gpu, tiny):
    s, eci = tiny["s"], tiny["eci"]
    occ = tiny["states"][7]
    e = tiny["energies"]
    w = (EMAX - EMIN) / NB
    occ_bins = tiny["g"] > 0
    hist = np.where(occ_bins, 100000, 0).astype(np.int32)
    cases = [
        dict(n_bins=NB, e_min=e.max() + 2 * w, e_max=e.max() + (2 + NB) * w, f=1.0, f_min=0.0, flatness=2.0,
             cur_bin=3, is_first=False),
        dict(n_bins=NB, e_min=e.min() - (2 + NB) * w, e_max=e.min() - 2 * w, f=1.0, f_min=0.0, flatness=2.0,
             cur_bin=3, is_first=False),
        dict(n_bins=NB, e_min=e.min() - (2 + NB) * w, e_max=e.min() - 2 * w, f=1.0, f_min=0.0, flatness=2.0,
             cur_bin=3, is_first=True),
        dict(n_bins=30, e_min=e.min() + 0.005, e_max=e.min() + 0.305, f=0.0, f_min=-1.0, flatness=2.0),
        dict(n_bins=NB, e_min=EMIN, e_max=EMAX, f=1.0, f_min=0.6, flatness=0.8, hist=hist,
             known=occ_bins.astype(np.uint8)),
        dict(n_bins=NB, e_min=EMIN, e_max=EMAX, f=0.0, f_min=-1.0, flatness=1.0, hist=hist,
             known=occ_bins.astype(np.uint8)),
    ]
    for k, c in enumerate(cases):
        print("range/flatness case", k, flush=True)
        _gpu_vs_oracle(s, eci, occ, c, 5, 700)


@pytest.mark.gpu
def test_gpu_walkers_of_several_models_in_one_batch(gpu):
    """Walkers on three models (table and closed-form lattices, two and
    three species) in one batch: each equals its own oracle walker; a chain
    without a walker is left alone."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    setups = [(FCCSettings(6, ["Al", "Mg"], {2: 3, 3: 2}), 2), (FCCSettings(5, ["Al", "Mg", "Si"], {2: 2, 3: 1}), 3),
              (FCCSettings((2, 2, 3), ["Al", "Mg"], {2: 2, 3: 1}), 2), (FCCSettings(6, ["Al", "Mg"], {2: 3, 3: 2}), 2)]
    models, chains, refs, wls = [], [], [], []
    n_checks, check_every = 6, 1000
    for c, (s, ns) in enumerate(setups):
        eci = seeded_ecis(s, 40 + c, 0.01)
        occ = random_occupation(s.n_sites, ns, 20 + c)
        names = s.cf_names()
        ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
        e0 = ora.get_energy()
        models.append(CEModel(s, eci))
        chains.append(dict(kind=CANONICAL, temperature=800.0, occupation=occ, seed=99, uid=c))
        args = dict(_centred(e0, 0.3, 30 + 5 * c), f=1.0, f_min=1e-3, flatness=0.6, cur_bin=15)
        if c < 3:
            wl = OracleWL(**args)
            bits = wl_run(ora, wl, 99, c, n_checks, check_every)
            refs.append((ora, wl, bits))
        wls.append(args)
    batch = ChainBatch(models, chains)
    for c in range(3):
        a = dict(wls[c])
        batch.wl_set(c, a.pop("n_bins"), a.pop("e_min"), a.pop("e_max"), a.pop("f"),
                     current_bin=a.pop("cur_bin"), **a)
    occ3 = batch.occupation(3).copy()
    log, ms = batch.wl_run(n_checks, check_every, log=True, timed=True)
    assert ms > 0.0
    for c, (ora, wl, bits) in enumerate(refs):
        got = batch.wl_get(c)
        n = int(wl.state.n_moves)
        assert got["n_moves"] == n
        assert np.array_equal(_bits(log[c], n), _bits(bits, n)), c
        assert np.array_equal(got["hist"], wl.hist) and np.array_equal(got["logdos"], wl.logdos)
        assert np.array_equal(batch.occupation(c), ora.get_occupation())
    assert np.array_equal(batch.occupation(3), occ3)
    assert not log[3].any()


@pytest.mark.gpu
def test_gpu_walker_resumes_across_launches(gpu):
    """Two launches of 3 checks equal one launch of 6 on the oracle."""
    from cemc_amd.batch import CANONICAL, ChainBatch
    from cemc_amd.model import CEModel
    s = FCCSettings(6, ["Al", "Mg"], {2: 3, 3: 2})
    eci = seeded_ecis(s, 8, 0.01)
    occ = random_occupation(s.n_sites, 2, 8)
    names = s.cf_names()
    ora = OracleUpdater(OracleModel(s, names), occ, brute_force_cf(s, occ, names), eci)
    e0 = ora.get_energy()
    args = dict(_centred(e0, 0.3, 40), f=1.0, f_min=1e-3, flatness=0.6, cur_bin=20)
    wl = OracleWL(**args)
    wl_run(ora, wl, 3, 4, 6, 1000)
    batch = ChainBatch([CEModel(s, eci)], [dict(kind=CANONICAL, temperature=800.0, occupation=occ, seed=3, uid=4)])
    a = dict(args)
    batch.wl_set(0, a.pop("n_bins"), a.pop("e_min"), a.pop("e_max"), a.pop("f"), current_bin=a.pop("cur_bin"), **a)
    batch.wl_run(3, 1000)
    batch.wl_run(3, 1000)
    got = batch.wl_get(0)
    assert got["n_moves"] == wl.state.n_moves
    assert np.array_equal(got["hist"], wl.hist) and np.array_equal(got["logdos"], wl.logdos)
    assert np.array_equal(batch.occupation(0), ora.get_occupation())


@pytest.mark.gpu
def test_python_driver_converges_to_exact_dos(gpu, tiny):
    """WangLandau on an atoms object with the CE calculator: eight walkers
    to f < 1e-4; their averaged ln g matches the enumerated one."""
    from cemc_amd.calculator import CE
    from cemc_amd.lattice import SimpleAtoms
    from cemc_amd.wang_landau import WangLandau
    s = tiny["s"]
    occ = tiny["states"][5]
    atoms = SimpleAtoms([s.species[v] for v in occ])
    CE(atoms, s, eci=dict(tiny["eci"]))
    wl = WangLandau(atoms, Nbins=NB, initial_f=1.0, flatness_criteria=0.8, fmin=1e-4, Emin=EMIN, Emax=EMAX,
                    check_convergence_every=2000, n_walkers=8, seed=4)
    assert wl.run_fast_sampler(maxsteps=400000)
    assert all(f == 2.0 ** -14 for f in wl.f_values)
    g = tiny["g"]
    known = g > 0
    est = wl.logdos()
    assert np.array_equal(~np.isnan(est), known)
    d = est[known] - np.log(g[known])
    d -= d.mean()
    assert np.abs(d).max() < 0.2, d
    wl.close()
