"""Continuous-batching request loop: scheduling logic on the CPU (with a
recording stand-in for the device batch), and on the GPU the property that a
job's result is the same alone and in the mix, whatever its admission step
and batch neighbours."""
import numpy as np
import pytest

from tis_amd import topology as T
from tis_amd.serve import Job, Server, estimate_bytes, poisson_stream


def _chain(n, seed):
    rng = np.random.default_rng(seed)
    stem = max(2, min(6, (n - 5) // 2))
    seq, pairs = T.hairpin_design(rng, n, stem, 3 if n < 30 else 5)
    s = T.build_rna_system([(seq, pairs)], tempk=310.0, ionic_strength=0.150, name=f"job{seed}")
    return s, T.helix_coords(s, rng)


class RecordingBatch:
    """Stands in for Engine in the CPU tests: every step adds (key + 1) * 1e-3
    to each coordinate of a live system, so a job's final coordinates tell
    how many steps it ran, and whether its state survived the batch changes.
    It records every state transfer by system index."""

    def __init__(self):
        self.appends, self.retires = [], []
        self.sys, self.keys, self.live = [], [], set()
        self.x, self.v, self.a, self.origin = [], [], [], []
        self.sets, self.gets = [], []
        self.max_live = 0

    def append(self, systems, keys=None):
        base = len(self.sys)
        for s, k in zip(systems, keys):
            self.sys.append(s)
            self.keys.append(k)
            for arr in (self.x, self.v, self.a):
                arr.append(np.zeros((s.nmp, 3)))
            self.origin.append(None)
        slots = list(range(base, len(self.sys)))
        self.live.update(slots)
        self.appends.append(list(keys))
        return slots

    def retire(self, slots):
        assert all(k in self.live for k in slots)
        self.live.difference_update(slots)
        self.retires.append([self.keys[k] for k in slots])

    def set_state(self, k, which, arr):
        assert k in self.live
        self.sets.append((k, which))
        [self.x, self.v, self.a][which][k] = np.array(arr)

    def get_state(self, k, which):
        assert k in self.live
        self.gets.append((k, which))
        return np.array([self.x, self.v, self.a][which][k])

    def system_origin(self, k, origin, init_accel):
        self.origin[k] = (origin, init_accel)

    def langevin_init(self, **kw):
        assert kw["init_accel"] is None
        self.step0 = kw["step0"]

    def device_bytes(self):
        return sum(estimate_bytes(self.sys[k]) for k in self.live)

    def transfer_bytes(self):
        return {"h2d_topology": 0, "h2d_state": 0, "d2h_state": 0, "d2d_state": 0}

    def arena_stats(self):
        return {"layouts": 0}

    def langevin_steps(self, n, mode):
        self.max_live = max(self.max_live, len(self.live))
        for k in self.live:
            self.x[k] = self.x[k] + n * (self.keys[k] + 1) * 1e-3


def _jobs(nsteps, arrivals, sizes):
    out = []
    for k, (ns, t, n) in enumerate(zip(nsteps, arrivals, sizes)):
        s, x = _chain(n, 100 + k)
        out.append(Job(key=k, system=s, x0=x, nsteps=ns, arrival=t))
    return out


def test_schedule_step_clock():
    jobs = _jobs([60, 120, 40, 100, 80], [0, 0, 20, 40, 45], [24, 40, 31, 64, 20])
    rb = RecordingBatch()
    srv = Server(round_steps=20, clock="steps", engine=rb)
    st = srv.run(jobs)
    assert st.jobs == 5 and len(srv.done) == 5
    for j in jobs:
        assert j.steps_done == j.nsteps
        np.testing.assert_allclose(j.x, j.x0 + j.nsteps * (j.key + 1) * 1e-3, rtol=0, atol=1e-12)
        # admitted at the first round boundary at or after arrival
        assert j.origin == 20 * int(np.ceil(j.arrival / 20))
        assert j.finished >= j.origin + j.nsteps
    # jobs 3 (arrival 40) and 4 (arrival 45, admitted at 60) join a running batch
    assert [3] in rb.appends and [4] in rb.appends
    assert st.max_inflight >= 3
    # each job's state crossed twice: set at admission, read at completion
    for k in range(5):
        assert sorted(w for i, w in rb.sets if i == k) == [0, 1, 2]
        assert sorted(w for i, w in rb.gets if i == k) == [0, 1, 2]


def test_budget_admission_fifo():
    jobs = _jobs([40, 40, 40, 40], [0, 0, 0, 0], [64, 64, 64, 64])
    est = estimate_bytes(jobs[0].system)
    rb = RecordingBatch()
    srv = Server(hbm_budget=2 * est + est // 2, round_steps=20, clock="steps", engine=rb)
    st = srv.run(jobs)
    assert st.jobs == 4
    assert rb.max_live == 2                              # never more than the budget allows
    assert [j.origin for j in jobs] == [0, 0, 40, 40]    # first come, first served
    for j in jobs:
        np.testing.assert_allclose(j.x, j.x0 + 40 * (j.key + 1) * 1e-3, rtol=0, atol=1e-12)


def test_submit_rejects():
    (j,) = _jobs([35], [0], [20])
    srv = Server(round_steps=20, clock="steps", engine=RecordingBatch())
    with pytest.raises(ValueError):
        srv.submit(j)
    (j,) = _jobs([40], [0], [20])
    srv = Server(hbm_budget=1000, round_steps=20, clock="steps", engine=RecordingBatch())
    with pytest.raises(ValueError):
        srv.submit(j)
    with pytest.raises(ValueError):
        Server(round_steps=25, rebuild_every=10)


def test_poisson_stream_and_latency_stats():
    rng = np.random.default_rng(3)
    sysx = [_chain(20, 7)]
    jobs = poisson_stream(rng, 200, 4.0, lambda k, r: Job(key=k, system=sysx[0][0], x0=sysx[0][1], nsteps=10))
    gaps = np.diff([0.0] + [j.arrival for j in jobs])
    assert abs(gaps.mean() - 0.25) < 0.05
    srv = Server(round_steps=10, clock="wall", engine=RecordingBatch())
    st = srv.run(jobs[:30])
    s = st.summary()
    assert s["jobs"] == 30 and s["p99_latency"] >= s["p50_latency"] >= 0.0


@pytest.mark.gpu
def test_alone_equals_mix_gpu():
    """Each job's final x, v, a in the mix are bitwise those of the job run
    alone, for two schedules (all admitted vs budget-limited queueing)."""
    from tis_amd.engine import Engine, MODE_GRAPH
    seed = 77
    spec = ([60, 120, 40, 100, 80], [0, 0, 20, 40, 45], [24, 40, 31, 64, 20])
    alone = []
    for j in _jobs(*spec):
        e = Engine([j.system], keys=[j.key], seed=seed)
        e.set_coords(0, j.x0)
        e.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
        e.langevin_steps(j.nsteps, MODE_GRAPH)
        alone.append([e.get_state(0, w) for w in range(3)])
        e.close()
    for budget in (1 << 34, None):
        jobs = _jobs(*spec)
        if budget is None:
            budget = 2 * max(estimate_bytes(j.system) for j in jobs)
        srv = Server(hbm_budget=budget, round_steps=20, clock="steps", seed=seed)
        st = srv.run(jobs)
        for j, ref in zip(jobs, alone):
            assert j.steps_done == j.nsteps
            for got, want in zip((j.x, j.v, j.a), ref):
                assert np.array_equal(got, want), f"job {j.key} differs (budget {budget})"
        # state crossed the bus only at admission and completion: x, v, a of
        # each job once each way (continuing jobs never left the device)
        state = sum(72 * j.system.nmp for j in jobs)
        assert st.h2d_state == state and st.d2h_state == state
        print(f"   budget {budget}: layouts {st.layouts}, h2d/admission "
              f"{st.summary()['h2d_bytes_per_admission']:.0f} B, carried on device {st.d2d_state} B", flush=True)
        srv._eng.close()


@pytest.mark.gpu
def test_alone_equals_mix_long_jobs_gpu():
    """Serving-size jobs: a 1000-nt and a 1200-nt job (2,999 and 3,599 beads)
    join a batch of short ones at different rounds; each job's final x, v,
    a equal its run alone, bitwise, and only admissions and completions
    move state across the bus."""
    from tis_amd.engine import Engine, MODE_GRAPH
    seed = 91
    spec = ([60, 80, 40, 60], [0, 20, 20, 40], [1000, 40, 1200, 64])
    alone = []
    for j in _jobs(*spec):
        print(f"[long jobs] alone: key {j.key}, {j.system.nmp} beads", flush=True)
        e = Engine([j.system], keys=[j.key], seed=seed)
        e.set_coords(0, j.x0)
        e.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
        e.langevin_steps(j.nsteps, MODE_GRAPH)
        alone.append([e.get_state(0, w) for w in range(3)])
        e.close()
    jobs = _jobs(*spec)
    srv = Server(round_steps=20, clock="steps", seed=seed)
    st = srv.run(jobs)
    for j, ref in zip(jobs, alone):
        for got, want in zip((j.x, j.v, j.a), ref):
            assert np.array_equal(got, want), f"job {j.key} differs"
    state = sum(72 * j.system.nmp for j in jobs)
    assert st.h2d_state == state and st.d2h_state == state
    srv._eng.close()
