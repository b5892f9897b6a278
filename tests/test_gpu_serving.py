"""GPU tests of the job server: every job's result is the same whether it
runs alone or in the mix of a continuously batched stream, admission
respects the pool's memory, and slots and memory are reused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(job):
    from cemc_amd.batch import ChainBatch
    b = ChainBatch([job.model], [dict(kind=job.kind, temperature=job.temperature, occupation=job.occupation,
                                      seed=job.seed, uid=job.uid, chem_pot=job.chem_pot)])
    b.run(job.n_moves)
    return b.energies()[0], b.occupation(0), int(b.stats()["accepted"][0])


def test_jobs_alone_and_in_the_mix(gpu):
    from cemc_amd.serving import JobServer, synthetic_stream
    jobs = synthetic_stream(n_jobs=24, rate=400.0, seed=8, sizes=(6, 10), sweeps=(20, 120))
    # a small quantum forces many launches with jobs joining and leaving
    server = JobServer(max_chains=16, hbm_bytes=1 << 20, quantum=3.0e5)
    res = server.serve(jobs, realtime=True)
    print(res, flush=True)
    assert res["jobs"] == len(jobs)
    assert res["launches"] > 10
    for job in server.done:
        e, occ, acc = _alone(job)
        assert np.array_equal(job.final_occupation, occ), job.job_id
        assert job.accepted == acc
        assert abs(job.energy - e) <= 1e-10 * job.model.n_sites
        assert job.finished >= job.admitted >= job.arrival - 1e-9


def test_admission_by_footprint(gpu):
    from cemc_amd._capi import PoolFull
    from cemc_amd.batch import CANONICAL, ChainPool
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    st = FCCSettings(8, ["A", "B"], {2: 2})
    model = CEModel(st, seeded_ecis(st, 1, 0.02))
    occ_b, trk = ChainPool.footprint(model, CANONICAL)
    assert occ_b == 512 and trk == 512
    pool = ChainPool(max_chains=8, max_models=2, occ_bytes=2 * 512, tracker_entries=8 * 512)
    chain = dict(kind=CANONICAL, temperature=600.0, occupation=random_occupation(st.n_sites, 2, 3), seed=1, uid=0)
    a = pool.add(model, chain)
    b = pool.add(model, dict(chain, uid=1))
    with pytest.raises(PoolFull):
        pool.add(model, dict(chain, uid=2))
    pool.remove(a)
    c = pool.add(model, dict(chain, uid=2))
    assert c == a
    nm = np.zeros(8, dtype=np.int64)
    nm[b] = 1000
    nm[c] = 500
    pool.run(nm)
    st_ = pool.stats()
    assert st_["moves"][b] == 1000 and st_["moves"][c] == 500


def test_time_sliced_jobs_alone_and_in_the_mix(gpu):
    """Time-sliced lanes: tiny slices cut every job's stream into many
    launches; each job still ends exactly as it does alone."""
    from cemc_amd.serving import JobServer, synthetic_stream
    jobs = synthetic_stream(n_jobs=20, rate=400.0, seed=9, sizes=(6, 12), sweeps=(20, 80))
    server = JobServer(max_chains=16, hbm_bytes=1 << 20, slice_ms=0.3)
    res = server.serve(jobs, realtime=True)
    print(res, flush=True)
    assert res["jobs"] == len(jobs)
    assert res["launches"] > 20
    assert res["moves"] == sum(j.n_moves for j in jobs)
    for job in server.done:
        e, occ, acc = _alone(job)
        assert np.array_equal(job.final_occupation, occ), job.job_id
        assert job.accepted == acc
        assert abs(job.energy - e) <= 1e-10 * job.model.n_sites


@pytest.mark.parametrize("shape", [(1, 1), (1, 8), (4, 8), (4, 32)])
def test_sliced_runs_continue_each_chains_stream(gpu, shape):
    """A batch advanced by short time slices until every chain is done ends
    bitwise where one uninterrupted run ends: lattices, energies, moves and
    accepted moves; the sample sums agree to rounding (each launch adds its
    partial sums to the chain's)."""
    from cemc_amd.batch import ChainBatch
    from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
    from cemc_amd.model import CEModel
    st = FCCSettings(14, ["A", "B"], {2: 4, 3: 2})
    model = CEModel(st, seeded_ecis(st, 5, 0.02))
    chains = [dict(kind=c % 2, temperature=400.0 + 300.0 * c, occupation=random_occupation(st.n_sites, 2, 40 + c),
                   seed=6, uid=c, chem_pot={"c1_0": 0.01}) for c in range(4)]
    n = 40000
    out = []
    for sliced in (False, True):
        b = ChainBatch([model] * 4, chains)
        b.set_moves_per_round(shape[1])
        b.set_waves(shape[0] if shape[1] <= 8 else 0)
        if sliced:
            launches = 0
            while True:
                left = n - b.stats()["moves"]
                if left.max() == 0:
                    break
                b.run_sliced(left, 0.2)
                launches += 1
            assert launches > 3
        else:
            b.run(n)
        out.append(b)
    a, b = out
    print("shape", shape, "launch shape", b.launch_shape(0), flush=True)
    assert np.array_equal(a.energies(), b.energies())
    for c in range(4):
        assert np.array_equal(a.occupation(c), b.occupation(c))
    sa, sb = a.stats(), b.stats()
    for key in ("moves", "accepted"):
        assert np.array_equal(sa[key], sb[key])
    for key in ("sum_e", "sum_e2"):
        assert np.allclose(sa[key], sb[key], rtol=1e-12, atol=1e-9)
