"""Continuous-batching request loop for DT15 Langevin jobs on one GPU.

A job is one system (a chain, or any molecule set) with starting coordinates
and a number of Langevin steps.  Jobs arrive over time; the server runs
rounds of at most ``round_steps`` steps over every admitted job in one batch
(one hipGraph-captured launch sequence per step), and between rounds admits
queued jobs as HBM budget frees up and retires finished ones.  A job keeps
its state (coordinates, velocities, accelerations) and its own step count
across the re-batching: its random streams are keyed by (seed, job key,
job-local step, bead), so its trajectory is bitwise the one it would follow
alone on the GPU.

The reference runs one job per process (mloop_simulator.F90 over its
replicas); this loop is what replaces many such processes on one device.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine, MODE_GRAPH
from .sysdesc import System

# bytes the library allocates per bead (positions x2, v, a, f, Langevin
# constants, masses, radii, charges, types, unit, charge index, incidence
# offset, epsilon, list count) plus the excluded-volume list at its initial
# capacity of 64 entries
_BEAD_BYTES = 48 + 72 + 32 + 32 + 3 + 8 + 4 + 4 + 4 + 64 * 4
_PERIODIC_BEAD_BYTES = 48


def estimate_bytes(s: System) -> int:
    """Device bytes the library allocates for system s (requested sizes).

    Per bead, per term, per incidence entry (4-byte entry + 24-byte force
    slot + 4-byte position), per charge (index, DH force, list bit matrix
    row) and per unit pair, as laid out by tis_system_upload."""
    n = s.nmp
    nbd, nba, nst = len(s.bd_nat), len(s.ba_nat), len(s.st_mp)
    ntst, nhb, nsite = len(s.tst_mp), len(s.hb_mp), len(s.hbsite_valence)
    periodic = bool(s.params.periodic)
    nq = int(np.count_nonzero(np.abs(s.charge) >= 1e-6)) if s.params.ele_form >= 0 else 0
    ninc = 2 * nbd + 3 * nba + 7 * nst + 6 * ntst + 6 * nhb
    b = n * (_BEAD_BYTES + (_PERIODIC_BEAD_BYTES if periodic else 0))
    b += ninc * 32
    b += nbd * 24 + nba * 28 + nst * 89 + ntst * 148 + nhb * 336 + nsite * 8
    b += nq * (4 + 24) + nq * ((nq + 31) // 32) * 4 + (nq * nq if periodic else 0)
    b += s.nunit * s.nunit * 8
    b += ((n + 63) // 64 + (nq + 63) // 64) * 56
    b += 512  # DSys, energy slots, indices
    return int(b)


@dataclass(eq=False)
class Job:
    """One simulation request."""
    key: int                       # identity in the random streams
    system: System
    x0: np.ndarray                 # (N, 3) starting coordinates
    nsteps: int                    # Langevin steps, a multiple of the rebuild interval
    arrival: float = 0.0           # seconds, or global steps with Server(clock="steps")
    v0: np.ndarray | None = None
    # filled by the server
    steps_done: int = 0
    origin: int = -1               # global step at admission
    admitted: float = math.nan
    finished: float = math.nan
    x: np.ndarray | None = None
    v: np.ndarray | None = None
    a: np.ndarray | None = None
    bytes: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeStats:
    jobs: int = 0
    rounds: int = 0
    rebatches: int = 0
    bead_steps: int = 0
    busy_s: float = 0.0            # wall time in rounds and re-batching
    span_s: float = 0.0            # first arrival to last completion (clock units)
    max_inflight: int = 0
    max_bytes: int = 0
    latencies: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.asarray(self.latencies, dtype=float)
        out = {"jobs": self.jobs, "rounds": self.rounds, "rebatches": self.rebatches,
               "bead_steps": self.bead_steps, "busy_s": self.busy_s, "span": self.span_s,
               "max_inflight": self.max_inflight, "max_device_bytes": self.max_bytes,
               "bead_steps_per_s_busy": self.bead_steps / self.busy_s if self.busy_s > 0 else None}
        if lat.size:
            out["p50_latency"] = float(np.percentile(lat, 50))
            out["p99_latency"] = float(np.percentile(lat, 99))
            out["mean_latency"] = float(lat.mean())
            if self.span_s > 0:
                out["mean_inflight"] = float(lat.sum() / self.span_s)  # Little's law
        return out


class Server:
    """Admit, batch and retire Langevin jobs on one device.

    hbm_budget: bytes the batch may hold on the device; a queued job is
    admitted (first come, first served) when the held bytes plus its
    estimate fit.  round_steps: steps per round (a multiple of
    rebuild_every), the longest a new arrival waits for admission.
    clock: "wall" (seconds; idle gaps before the next arrival are skipped,
    not slept) or "steps" (the global step count: deterministic schedules).
    """

    def __init__(self, hbm_budget: int = 32 << 30, round_steps: int = 1000, tstep: float = 0.05,
                 rebuild_every: int = 10, seed: int = 1, device: int = 0, clock: str = "wall",
                 max_jobs: int = 4096, engine=None):
        if round_steps % rebuild_every:
            raise ValueError("round_steps must be a multiple of rebuild_every")
        if clock not in ("wall", "steps"):
            raise ValueError("clock is 'wall' or 'steps'")
        self.budget = int(hbm_budget)
        self.round_steps = int(round_steps)
        self.tstep = float(tstep)
        self.rebuild_every = int(rebuild_every)
        self.seed = int(seed)
        self.device = device
        self.clock = clock
        self.max_jobs = int(max_jobs)
        self._eng = engine
        self.queue: list[Job] = []
        self.active: list[Job] = []
        self.batch: list[Job] = []     # jobs in the device batch, in slot order
        self._pulled = False           # batch state copied to the jobs since the last round
        self.done: list[Job] = []
        self.gstep = 0
        self.held_bytes = 0
        self.stats = ServeStats()
        self._t0 = None
        self._skip = 0.0

    # ---- clock
    def now(self) -> float:
        if self.clock == "steps":
            return float(self.gstep)
        return time.perf_counter() - self._t0 + self._skip

    # ---- requests
    def submit(self, job: Job):
        if job.nsteps <= 0 or job.nsteps % self.rebuild_every:
            raise ValueError(f"job {job.key}: nsteps must be a positive multiple of {self.rebuild_every}")
        if job.x0.shape != (job.system.nmp, 3):
            raise ValueError(f"job {job.key}: x0 must be (N, 3)")
        job.bytes = estimate_bytes(job.system)
        if job.bytes > self.budget:
            raise ValueError(f"job {job.key} needs {job.bytes} B, over the HBM budget {self.budget} B")
        self.queue.append(job)

    def _engine(self, systems, keys):
        if self._eng is None:
            self._eng = Engine(systems, keys=keys, seed=self.seed, device=self.device)
        else:
            self._eng.upload(systems, keys=keys, seed=self.seed)
        return self._eng

    def _admit(self, now) -> list[Job]:
        new = []
        held = sum(j.bytes for j in self.active)
        for job in list(self.queue):
            if job.arrival > now:
                break
            if len(self.active) + len(new) >= self.max_jobs or held + job.bytes > self.budget:
                break  # first come, first served: no overtaking
            held += job.bytes
            new.append(job)
            self.queue.remove(job)
        return new

    def _rebatch(self, new: list[Job]):
        """Upload the active set; continuing jobs carry their state over."""
        for j in new:
            j.origin = self.gstep
            j.admitted = self.now()
            n = j.system.nmp
            j.x = np.array(j.x0, dtype=np.float64)
            j.v = np.zeros((n, 3)) if j.v0 is None else np.array(j.v0, dtype=np.float64)
            j.a = np.zeros((n, 3))
        self.active.extend(new)
        self.batch = list(self.active)
        self._pulled = False
        if not self.batch:
            return
        eng = self._engine([j.system for j in self.batch], [j.key for j in self.batch])
        for k, j in enumerate(self.batch):
            eng.set_state(k, 0, j.x)
            eng.set_state(k, 1, j.v)
            eng.set_state(k, 2, j.a)
            eng.system_origin(k, j.origin, j.steps_done == 0)
        eng.langevin_init(tstep=self.tstep, rebuild_every=self.rebuild_every, step0=self.gstep, init_accel=None)
        self.held_bytes = eng.device_bytes()
        self.stats.rebatches += 1
        self.stats.max_bytes = max(self.stats.max_bytes, self.held_bytes)
        self.stats.max_inflight = max(self.stats.max_inflight, len(self.active))

    def _pull_state(self):
        if self._pulled:
            return
        eng = self._eng
        for k, j in enumerate(self.batch):
            j.x = eng.get_state(k, 0)
            j.v = eng.get_state(k, 1)
            j.a = eng.get_state(k, 2)
        self._pulled = True

    def run(self, jobs: list[Job] | None = None) -> ServeStats:
        """Serve until every submitted job has finished."""
        for j in sorted(jobs or [], key=lambda j: j.arrival):
            self.submit(j)
        self.queue.sort(key=lambda j: j.arrival)
        self._t0 = time.perf_counter()
        first = self.queue[0].arrival if self.queue else 0.0
        while self.queue or self.active:
            now = self.now()
            if not self.active and self.queue and self.queue[0].arrival > now:
                if self.clock == "wall":
                    self._skip += self.queue[0].arrival - now   # idle: jump to the next arrival
                else:
                    # step clock with nothing to run: advance whole rounds
                    gap = self.queue[0].arrival - now
                    self.gstep += int(math.ceil(gap / self.round_steps)) * self.round_steps
                continue
            tb = time.perf_counter()
            new = self._admit(self.now())
            if new or [id(j) for j in self.batch] != [id(j) for j in self.active]:
                if self.batch:
                    self._pull_state()
                self._rebatch(new)
            if not self.active:
                self.stats.busy_s += time.perf_counter() - tb
                continue
            nrun = min(self.round_steps, min(j.nsteps - j.steps_done for j in self.active))
            self._eng.langevin_steps(nrun, MODE_GRAPH)
            self._pulled = False
            self.gstep += nrun
            self.stats.rounds += 1
            self.stats.bead_steps += nrun * sum(j.system.nmp for j in self.active)
            for j in self.active:
                j.steps_done += nrun
            finished = [j for j in self.active if j.steps_done >= j.nsteps]
            if finished:
                self._pull_state()
                t = self.now()
                for j in finished:
                    j.finished = t
                    self.stats.latencies.append(j.latency)
                    self.done.append(j)
                self.active = [j for j in self.active if j.steps_done < j.nsteps]
                self.stats.jobs += len(finished)
            self.stats.busy_s += time.perf_counter() - tb
        if self.done:
            self.stats.span_s = max(j.finished for j in self.done) - first
        return self.stats


def poisson_stream(rng: np.random.Generator, njobs: int, rate: float, make_job) -> list[Job]:
    """njobs jobs with exponential inter-arrival times (mean 1/rate)."""
    t = 0.0
    out = []
    for k in range(njobs):
        t += float(rng.exponential(1.0 / rate))
        job = make_job(k, rng)
        job.arrival = t
        out.append(job)
    return out
