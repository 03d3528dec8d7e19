"""Continuous-batching request loop for DT15 Langevin jobs on one GPU.

A job is one system (a chain, or any molecule set) with starting coordinates
and a number of Langevin steps.  Jobs arrive over time; the server runs
rounds of at most ``round_steps`` steps over every admitted job in one batch
(one hipGraph-captured launch sequence per step), and between rounds admits
queued jobs as HBM budget frees up and retires finished ones.

The batch lives in the library's arena (tis_system_append /
tis_system_retire): an admission uploads only the new jobs' topology and
starting state, a completion downloads only the finished jobs' state, and
the jobs that continue never leave the device.  A job keeps its own step
count: its random streams are keyed by (seed, job key, job-local step,
bead), so its trajectory is bitwise the one it would follow alone.

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
    slot: int = -1                 # system index in the device batch
    fresh: bool = True             # its accelerations are still to be initialised

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
    admitted: int = 0
    # payload bytes (tis_transfer_bytes) of admissions and completions
    h2d_topology: int = 0
    h2d_state: int = 0
    d2h_state: int = 0
    d2d_state: int = 0
    layouts: int = 0

    def summary(self) -> dict:
        lat = np.asarray(self.latencies, dtype=float)
        out = {"jobs": self.jobs, "rounds": self.rounds, "rebatches": self.rebatches,
               "bead_steps": self.bead_steps, "busy_s": self.busy_s, "span": self.span_s,
               "max_inflight": self.max_inflight, "max_device_bytes": self.max_bytes,
               "bead_steps_per_s_busy": self.bead_steps / self.busy_s if self.busy_s > 0 else None,
               "h2d_bytes_per_admission": (self.h2d_topology + self.h2d_state) / max(self.admitted, 1),
               "h2d_state_bytes": self.h2d_state, "d2h_state_bytes": self.d2h_state,
               "d2h_bytes_per_completion": self.d2h_state / max(self.jobs, 1),
               "d2d_state_bytes": self.d2d_state, "layouts": self.layouts}
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

    def _append(self, new: list[Job]):
        """Append the new jobs to the device batch: their topology and
        starting state are the only data that cross to the device."""
        for j in new:
            j.origin = self.gstep
            j.admitted = self.now()
        systems, keys = [j.system for j in new], [j.key for j in new]
        if self._eng is None:
            self._eng = Engine(systems, keys=keys, seed=self.seed, device=self.device)
            slots = list(range(len(new)))
        else:
            slots = self._eng.append(systems, keys=keys)
        eng = self._eng
        for j, h in zip(new, slots):
            n = j.system.nmp
            j.slot = h
            j.fresh = True
            eng.set_state(h, 0, j.x0)
            eng.set_state(h, 1, np.zeros((n, 3)) if j.v0 is None else j.v0)
            eng.set_state(h, 2, np.zeros((n, 3)))
            eng.system_origin(h, j.origin, True)
        self.active.extend(new)
        self.stats.admitted += len(new)

    def _retire(self, finished: list[Job]):
        """Download the finished jobs' state and drop them from the batch."""
        eng = self._eng
        for j in finished:
            j.x = eng.get_state(j.slot, 0)
            j.v = eng.get_state(j.slot, 1)
            j.a = eng.get_state(j.slot, 2)
        eng.retire([j.slot for j in finished])

    def _restart(self):
        """Langevin set-up after a change of the batch: new jobs get new
        accelerations, continuing ones keep theirs (mode 3)."""
        eng = self._eng
        for j in self.active:
            if j.fresh and j.steps_done > 0:
                eng.system_origin(j.slot, j.origin, False)
                j.fresh = False
        eng.langevin_init(tstep=self.tstep, rebuild_every=self.rebuild_every, step0=self.gstep, init_accel=None)
        self.held_bytes = eng.device_bytes()
        self.stats.rebatches += 1
        self.stats.max_bytes = max(self.stats.max_bytes, self.held_bytes)
        self.stats.max_inflight = max(self.stats.max_inflight, len(self.active))

    def _count_bytes(self, before: dict):
        after = self._eng.transfer_bytes()
        self.stats.h2d_topology += after["h2d_topology"] - before.get("h2d_topology", 0)
        self.stats.h2d_state += after["h2d_state"] - before.get("h2d_state", 0)
        self.stats.d2h_state += after["d2h_state"] - before.get("d2h_state", 0)
        self.stats.d2d_state += after["d2d_state"] - before.get("d2d_state", 0)
        self.stats.layouts = self._eng.arena_stats()["layouts"]

    def run(self, jobs: list[Job] | None = None) -> ServeStats:
        """Serve until every submitted job has finished."""
        for j in sorted(jobs or [], key=lambda j: j.arrival):
            self.submit(j)
        self.queue.sort(key=lambda j: j.arrival)
        self._t0 = time.perf_counter()
        first = self.queue[0].arrival if self.queue else 0.0
        dirty = False
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
            if new:
                before = self._eng.transfer_bytes() if self._eng is not None else {}
                self._append(new)
                self._count_bytes(before)
                dirty = True
            if not self.active:
                self.stats.busy_s += time.perf_counter() - tb
                continue
            if dirty:
                self._restart()
                dirty = False
            nrun = min(self.round_steps, min(j.nsteps - j.steps_done for j in self.active))
            self._eng.langevin_steps(nrun, MODE_GRAPH)
            self.gstep += nrun
            self.stats.rounds += 1
            self.stats.bead_steps += nrun * sum(j.system.nmp for j in self.active)
            for j in self.active:
                j.steps_done += nrun
            finished = [j for j in self.active if j.steps_done >= j.nsteps]
            if finished:
                before = self._eng.transfer_bytes()
                self._retire(finished)
                self._count_bytes(before)
                dirty = True
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
