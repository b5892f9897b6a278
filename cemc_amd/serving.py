"""Job server: Monte Carlo jobs admitted onto one GPU as they arrive, with
continuous batching.

Each job is one Markov chain (a lattice, a move kind, a temperature, a number
of sweeps).  The server keeps a ``ChainPool``: a job is admitted as soon as a
chain slot and device memory for its lattice are free (admission by HBM
footprint), its lattice then stays resident on the GPU until the job ends.
The server loop runs quanta: one launch advances every admitted job by a
share of its remaining moves sized by the job's cost per move, so jobs that
arrive while others are running join the next launch, and finished jobs
leave and free their memory.  A job's moves are the same moves it makes when
run alone (each chain consumes its own Philox stream, counter = move index).

Reported per job: arrival, admission and completion times (wall clock of the
serving loop); the stream's latency percentiles and throughput.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from .batch import CANONICAL, SGC, ChainPool


@dataclass
class Job:
    job_id: int
    arrival: float                  # seconds after the start of the stream
    model: object                   # CEModel
    kind: int
    temperature: float
    sweeps: int
    occupation: np.ndarray
    seed: int = 0
    uid: int = 0
    chem_pot: dict = None
    # filled by the server
    admitted: float = None
    finished: float = None
    energy: float = None
    final_occupation: np.ndarray = None
    accepted: int = 0
    remaining: int = 0
    slot: int = -1
    extra: dict = field(default_factory=dict)

    @property
    def n_moves(self):
        return int(self.sweeps) * int(self.model.n_sites)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class JobServer:
    """Continuous-batching server over a ChainPool.

    :param max_chains: chain slots (jobs resident at once)
    :param hbm_bytes: device memory for resident lattices (occupation bytes
        plus 2 bytes per site of canonical species trackers)
    :param quantum: work per job per launch, in member gathers (one flip of
        a site reads every non-changed member of its sub-clusters); a job
        makes quantum / (gathers per flip x flips per move) moves per launch
    """

    def __init__(self, max_chains=2048, max_models=64, hbm_bytes=256 << 20, quantum=2.0e7):
        occ_bytes = hbm_bytes // 3
        tracker_entries = (hbm_bytes - occ_bytes) // 2
        self.pool = ChainPool(max_chains, max_models, occ_bytes, tracker_entries)
        self.quantum = float(quantum)
        self.queue = deque()
        self.running = {}
        self.done = []
        self.launches = 0
        self.moves_done = 0

    def submit(self, job):
        job.remaining = job.n_moves
        self.queue.append(job)

    def _admit(self, now):
        while self.queue:
            job = self.queue[0]
            chain = dict(kind=job.kind, temperature=job.temperature, occupation=job.occupation, seed=job.seed,
                         uid=job.uid, chem_pot=job.chem_pot)
            try:
                slot = self.pool.add(job.model, chain)
            except _capi.PoolFull:
                break
            self.queue.popleft()
            job.slot = slot
            job.admitted = now
            self.running[slot] = job

    def _share(self, job):
        flips = 2 if job.kind == CANONICAL else 1
        return max(64, int(self.quantum / max(1, job.model.gathers_per_flip() * flips)))

    def step(self, clock, t0):
        """One launch over the admitted jobs; finished jobs leave the pool."""
        if not self.running:
            return 0
        nm = np.zeros(self.pool.n_chains, dtype=np.int64)
        for slot, job in self.running.items():
            nm[slot] = min(job.remaining, self._share(job))
        self.pool.run(nm)
        self.launches += 1
        self.moves_done += int(nm.sum())
        finished = [slot for slot, job in self.running.items() if job.remaining - nm[slot] == 0]
        for slot, job in self.running.items():
            job.remaining -= int(nm[slot])
        if finished:
            now = clock() - t0
            energies = self.pool.energies()
            acc = self.pool.stats()["accepted"]
            for slot in finished:
                job = self.running.pop(slot)
                job.finished = now
                job.energy = float(energies[slot])
                job.final_occupation = self.pool.occupation(slot)
                job.accepted = int(acc[slot])
                self.pool.remove(slot)
                self.done.append(job)
        return len(finished)

    def serve(self, jobs, clock=time.perf_counter, realtime=True, progress=None, progress_every=10.0):
        """Run a stream of jobs (sorted by arrival).  realtime: jobs become
        visible at their arrival times on the wall clock; otherwise all at
        once.  progress(msg) is called about every progress_every seconds."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        t0 = clock()
        next_report = progress_every
        while pending or self.queue or self.running:
            now = clock() - t0
            if progress is not None and now >= next_report:
                progress("serving t=%.0f s: %d done, %d running, %d queued, %d not arrived, %.3e moves"
                         % (now, len(self.done), len(self.running), len(self.queue), len(pending), self.moves_done))
                next_report = now + progress_every
            while pending and (not realtime or pending[0].arrival <= now):
                self.submit(pending.popleft())
            self._admit(now)
            if self.running:
                self.step(clock, t0)
            elif pending:
                time.sleep(min(max(pending[0].arrival - now, 0.0), 0.002))
        makespan = clock() - t0
        return summarize(self.done, makespan, self.moves_done, self.launches)


def summarize(done, makespan, moves, launches):
    lat = np.array([j.latency for j in done]) if done else np.zeros(1)
    return {"jobs": len(done), "makespan_s": makespan, "moves": int(moves), "launches": int(launches),
            "throughput_moves_per_s": moves / makespan if makespan > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max())}


def synthetic_stream(n_jobs=5000, rate=200.0, seed=4, sizes=(6, 24), sweeps=(100, 5000), model_cache=None):
    """The serving workload: Poisson arrivals at `rate` jobs/s; binary or
    ternary FCC (1/2 each) with side L uniform in sizes (L^3 sites), T uniform
    in 300-1500 K, canonical or SGC (1/2 each), sweeps uniform in `sweeps`.
    Binary jobs use pairs to the 4th shell and triplets to the 2nd; ternary
    jobs add triplets to the 3rd and quadruplets to the 2nd shell.  SGC
    chemical potentials are uniform in [-0.1, 0.1] eV per singlet."""
    from .lattice import FCCSettings, random_occupation, seeded_ecis
    from .model import CEModel
    cache = {} if model_cache is None else model_cache
    rng = np.random.Generator(np.random.Philox(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    jobs = []
    for i in range(n_jobs):
        L = int(rng.integers(sizes[0], sizes[1] + 1))
        ternary = bool(rng.integers(0, 2))
        kind = CANONICAL if rng.integers(0, 2) == 0 else SGC
        T = float(rng.uniform(300.0, 1500.0))
        n_sweeps = int(rng.integers(sweeps[0], sweeps[1] + 1))
        key = (L, ternary)
        if key not in cache:
            species = ["A", "B", "C"] if ternary else ["A", "B"]
            shells = {2: 4, 3: 3, 4: 2} if ternary else {2: 4, 3: 2}
            st = FCCSettings(L, species, shells)
            cache[key] = CEModel(st, seeded_ecis(st, 2027 if ternary else 2024, 0.01 if ternary else 0.02))
        model = cache[key]
        n_sp = 3 if ternary else 2
        occ = random_occupation(model.n_sites, n_sp, 90000 + i)
        mu = None
        if kind == SGC:
            mu = {"c1_%d" % d: float(rng.uniform(-0.1, 0.1)) for d in range(n_sp - 1)}
        jobs.append(Job(job_id=i, arrival=float(arrivals[i]), model=model, kind=kind, temperature=T,
                        sweeps=n_sweeps, occupation=occ, seed=77, uid=i, chem_pot=mu))
    return jobs
