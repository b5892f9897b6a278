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
