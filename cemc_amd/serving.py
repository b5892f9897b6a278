"""Job server: Monte Carlo jobs admitted onto one GPU as they arrive, with
continuous batching.

Each job is one Markov chain (a lattice, a move kind, a temperature, a number
of sweeps).  Jobs go to lanes by species count and move kind; each lane keeps
a ``ChainPool`` and runs on its own host thread, so the lanes' launches
overlap on the GPU.  A job is admitted as soon as a chain slot and device
memory for its lattice are free (admission by HBM footprint), its lattice
then stays resident on the GPU until the job ends.  A lane runs time-sliced
launches: every admitted job moves until it is done or the slice's device
time is used up (each chain checks the device clock every 64 moves), so no
job waits for a slower one; jobs that arrive while others are running join
the next launch, and finished jobs leave and free their memory.  A job's moves are the same moves it makes when
run alone (each chain consumes its own Philox stream, counter = move index).

Reported per job: arrival, admission and completion times (wall clock of the
serving loop); the stream's latency percentiles and throughput.
"""
from __future__ import annotations

import threading
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


class _Lane:
    """One pool of resident jobs and the host thread that launches it."""

    def __init__(self, key, max_chains, max_models, hbm_bytes, quantum, slice_ms):
        occ_bytes = hbm_bytes // 3
        tracker_entries = (hbm_bytes - occ_bytes) // 2
        self.key = key
        self.pool = ChainPool(max_chains, max_models, occ_bytes, tracker_entries)
        self.quantum = None if quantum is None else float(quantum)
        self.slice_ms = float(slice_ms)
        self.inbox = deque()    # arrived, not yet seen by the lane's thread
        self.queue = deque()    # waiting for a slot or memory
        self.running = {}
        self.done = []
        self.rejected = []
        self.launches = 0
        self.moves_done = 0
        self.closed = False     # no more jobs will arrive
        self.lock = threading.Lock()
        self.error = None

    def admit(self, now):
        while self.queue:
            job = self.queue[0]
            chain = dict(kind=job.kind, temperature=job.temperature, occupation=job.occupation, seed=job.seed,
                         uid=job.uid, chem_pot=job.chem_pot)
            try:
                slot = self.pool.add(job.model, chain)
            except _capi.PoolFull as exc:
                if self.running:
                    break  # wait for resident jobs to leave
                # nothing is resident, so the job can never fit: reject it
                # instead of waiting forever
                self.queue.popleft()
                job.extra["rejected"] = str(exc)
                self.rejected.append(job)
                continue
            self.queue.popleft()
            job.slot = slot
            job.admitted = now
            job.extra["first_move"] = 0
            self.running[slot] = job

    def share(self, job):
        flips = 2 if job.kind == CANONICAL else 1
        return max(64, int(self.quantum / max(1, job.model.gathers_per_flip() * flips)))

    def step(self, clock, t0):
        """One launch over the lane's resident jobs; finished jobs leave.
        Time slices (quantum None): every job may make all its remaining
        moves but stops after slice_ms of device time, so no job waits for a
        slower one; quanta: every job makes a share of its remaining moves."""
        nm = np.zeros(self.pool.n_chains, dtype=np.int64)
        slots = list(self.running)
        if self.quantum is None:
            for slot in slots:
                nm[slot] = self.running[slot].remaining
            self.pool.run_sliced(nm, self.slice_ms)
            done_now = self.pool.stats()["moves"]
            made = np.zeros_like(nm)
            for slot in slots:
                job = self.running[slot]
                made[slot] = (int(done_now[slot]) - job.extra["first_move"]) - (job.n_moves - job.remaining)
        else:
            for slot in slots:
                nm[slot] = min(self.running[slot].remaining, self.share(self.running[slot]))
            self.pool.run(nm)
            made = nm
        self.launches += 1
        self.moves_done += int(made.sum())
        finished = [slot for slot in slots if self.running[slot].remaining - made[slot] == 0]
        for slot in slots:
            self.running[slot].remaining -= int(made[slot])
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

    def serve(self, clock, t0):
        try:
            while True:
                with self.lock:
                    while self.inbox:
                        self.queue.append(self.inbox.popleft())
                    closed = self.closed
                self.admit(clock() - t0)
                if self.running:
                    self.step(clock, t0)
                elif not self.queue and closed:
                    with self.lock:
                        if not self.inbox:
                            return
                else:
                    time.sleep(0.0005)
        except BaseException as exc:  # reported by JobServer.serve
            self.error = exc


class JobServer:
    """Continuous-batching server.

    Jobs are split into lanes by (species count, move kind); each lane is a
    ``ChainPool`` served by its own host thread, so the launches of different
    lanes run side by side on the GPU and a lane never waits for another
    lane's slowest chain.  Within a lane each launch is a time slice: every
    resident job moves until it is done or the slice is used up, jobs that
    arrived meanwhile join the next launch, finished jobs leave and free
    their slot and memory.

    :param max_chains: chain slots per lane (jobs resident at once)
    :param hbm_bytes: device memory per lane for resident lattices
        (occupation bytes plus 2 bytes per site of canonical species
        trackers); a job is admitted when its footprint fits
    :param slice_ms: device time per launch: every resident job runs until
        it is done or its slice is used up (time-sliced launches)
    :param quantum: instead of time slices, work per job per launch in member
        gathers (one flip of a site reads every non-changed member of its
        sub-clusters): a job makes quantum / (gathers per flip x flips per
        move) moves per launch
    :param lanes: False puts every job in one lane (one pool, one thread)
    """

    def __init__(self, max_chains=2048, max_models=64, hbm_bytes=256 << 20, quantum=None, lanes=True,
                 slice_ms=20.0):
        self.max_chains = int(max_chains)
        self.max_models = int(max_models)
        self.hbm_bytes = int(hbm_bytes)
        self.quantum = quantum
        self.slice_ms = float(slice_ms)
        self.use_lanes = bool(lanes)
        self.lanes = {}
        self.done = []
        self.rejected = []
        self.launches = 0
        self.moves_done = 0

    def lane_key(self, job):
        return (job.model.n_species, job.kind) if self.use_lanes else ()

    def _lane(self, key):
        if key not in self.lanes:
            self.lanes[key] = _Lane(key, self.max_chains, self.max_models, self.hbm_bytes, self.quantum,
                                    self.slice_ms)
        return self.lanes[key]

    def submit(self, job):
        job.remaining = job.n_moves
        lane = self._lane(self.lane_key(job))
        with lane.lock:
            lane.inbox.append(job)

    def serve(self, jobs, clock=time.perf_counter, realtime=True, progress=None, progress_every=10.0):
        """Run a stream of jobs.  realtime: jobs are handed to their lanes at
        their arrival times on the wall clock; otherwise all at once.
        progress(msg) is called about every progress_every seconds."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        for job in pending:  # lanes (and their pools) exist before the clock starts
            self._lane(self.lane_key(job))
        t0 = clock()
        threads = [threading.Thread(target=lane.serve, args=(clock, t0), daemon=True)
                   for lane in self.lanes.values()]
        for th in threads:
            th.start()
        next_report = progress_every
        while pending:
            now = clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                self.submit(pending.popleft())
            if progress is not None and now >= next_report:
                self._report(progress, now, len(pending))
                next_report = now + progress_every
            if pending:
                time.sleep(min(max(pending[0].arrival - (clock() - t0), 0.0), 0.002))
        for lane in self.lanes.values():
            with lane.lock:
                lane.closed = True
        while any(th.is_alive() for th in threads):
            for th in threads:
                th.join(timeout=min(progress_every, 1.0))
            now = clock() - t0
            if progress is not None and now >= next_report:
                self._report(progress, now, 0)
                next_report = now + progress_every
        makespan = clock() - t0
        for lane in self.lanes.values():
            if lane.error is not None:
                raise lane.error
        self.done = sorted((j for lane in self.lanes.values() for j in lane.done), key=lambda j: j.job_id)
        self.rejected = [j for lane in self.lanes.values() for j in lane.rejected]
        self.launches = sum(lane.launches for lane in self.lanes.values())
        self.moves_done = sum(lane.moves_done for lane in self.lanes.values())
        return summarize(self.done, makespan, self.moves_done, self.launches)

    def _report(self, progress, now, not_arrived):
        lanes = list(self.lanes.values())
        progress("serving t=%.0f s: %d done, %d running, %d queued, %d not arrived, %.3e moves"
                 % (now, sum(len(la.done) for la in lanes), sum(len(la.running) for la in lanes),
                    sum(len(la.queue) + len(la.inbox) for la in lanes), not_arrived,
                    sum(la.moves_done for la in lanes)))


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
