"""Replica exchange: host logic on the CPU with a recording stand-in for the
device batch (closed-form acceptance for temperature-independent energies,
alternating pairs, velocity rescale), the host Philox against the oracle's,
and on the GPU the swapped-temperature energies against the oracle."""
import math

import numpy as np
import pytest

import oracle
from tis_amd import rng
from tis_amd import topology as T
from tis_amd.remd import ReplicaLadder
from tis_amd.sysdesc import E_TOTAL
from tis_amd.topology import BOLTZ_KCAL_MOL


def test_host_uniform_matches_oracle_stream():
    for args in [(0, 0, 0, 0, 0), (20261202, 5, 123456789012, 1, 7), (2 ** 40 + 3, 65535, 2 ** 33 + 5, 3, 99)]:
        assert rng.uniform(*args) == oracle.uniform(*args)
    assert rng.philox4x32([0, 0, 0, 0], [0, 0]) == oracle.philox([0, 0, 0, 0], [0, 0])


class FixedEnergies:
    """Stand-in batch: system k has energy E[k] at every temperature."""

    def __init__(self, energies, nmp=5):
        self.E = np.asarray(energies, float)
        self.T = [None] * len(energies)
        self.v = [np.ones((nmp, 3)) * (k + 1) for k in range(len(energies))]

    def label_define(self, label, isys, sysd):
        self.labels = getattr(self, "labels", {})
        self.labels[label] = sysd.tempk

    def label_apply(self, isys, labels, vscale=None):
        for n, (k, lab) in enumerate(zip(isys, labels)):
            self.T[k] = self.labels[lab]
            if vscale is not None:
                self.v[k] = self.v[k] * vscale[n]

    def energy_all(self):
        e = np.zeros((len(self.E), 25))
        e[:, E_TOTAL] = self.E
        return e

    def energy_systems(self, isys):
        return self.energy_all()[list(isys)]

    def get_state(self, k, which):
        assert which == 1
        return self.v[k].copy()

    def set_state(self, k, which, arr):
        assert which == 1
        self.v[k] = np.array(arr)


def _base():
    seq, pairs = T.config0_chain()
    return T.build_rna_system([(seq, pairs)])


def test_acceptance_matches_closed_form():
    temps = [300.0, 320.0, 345.0, 370.0]
    E = [-50.0, -80.0, -20.0, -20.5]
    eng = FixedEnergies(E)
    lad = ReplicaLadder(eng, [0, 1, 2, 3], _base(), temps, seed=7, key=3)
    assert eng.T == temps
    res = lad.attempt()
    assert [(a, b) for a, b, _, _ in res] == [(0, 1), (2, 3)]
    for n, (li, lj, delta, acc) in enumerate(res):
        want = (E[lj] - E[li]) * (1 / temps[li] - 1 / temps[lj]) / BOLTZ_KCAL_MOL
        assert abs(delta - want) <= 1e-12 * abs(want)
        u = rng.uniform(7, 3, 0, 0, n)
        assert acc == (want < 0 or u <= math.exp(-want))
    # pair (0,1): E1 < E0 -> delta < 0 -> swapped; replica 0 now at label 1, heated
    assert res[0][3] and lad.rep2lab[0] == 1 and lad.lab2rep[1] == 0
    assert np.allclose(eng.v[0], 1.0 * math.sqrt(320.0 / 300.0))
    assert np.allclose(eng.v[1], 2.0 / math.sqrt(320.0 / 300.0), rtol=1e-14)
    assert eng.T[0] == 320.0 and eng.T[1] == 300.0
    res2 = lad.attempt()
    assert [(a, b) for a, b, _, _ in res2] == [(1, 2)]


@pytest.mark.gpu
def test_swapped_temperature_energies_gpu():
    """Energy of each replica at its partner's temperature equals the
    oracle's energy of the same coordinates with that temperature's
    parameters (energy_allrep.F90:150-216)."""
    from tis_amd.engine import Engine, MODE_GRAPH
    base = _base()
    temps = [300.0, 330.0, 360.0, 390.0]
    systems = []
    for t in temps:
        s = T.build_rna_system([T.config0_chain()])
        T.set_temperature(s, t, 0.150)
        systems.append(s)
    eng = Engine(systems, keys=[0, 1, 2, 3], seed=5)
    rng_ = np.random.default_rng(2)
    for k, s in enumerate(systems):
        eng.set_coords(k, T.helix_coords(s, rng_))
    lad = ReplicaLadder(eng, [0, 1, 2, 3], base, temps, seed=9, key=1)
    eng.langevin_init(tstep=0.05, rebuild_every=10)
    eng.langevin_steps(100, MODE_GRAPH)
    eng.force_all()
    own, part = lad.energies()
    for r in range(4):
        x = eng.coords(r)
        for lab, want in ((lad.rep2lab[r], own[r]), (lad.partner(lad.rep2lab[r]), part[r])):
            if lab < 0:
                continue
            o = oracle.Oracle(lad.label_sys[lab])
            o.neighbor(x)
            o.force(x, seed=5, sys_id=r, step=100)
            e = o.energy(x, eng.get_state(r, 1))[E_TOTAL]
            assert abs(e - want) <= 1e-7 * abs(e), (r, lab, e, want)
    v_before = [eng.get_state(r, 1) for r in range(4)]
    res = lad.attempt()
    for li, lj, delta, acc in res:
        ri, rj = lad.lab2rep[lj if acc else li], lad.lab2rep[li if acc else lj]
        if acc:
            f = math.sqrt(temps[lj] / temps[li])
            assert np.allclose(eng.get_state(ri, 1), v_before[ri] * f, rtol=1e-15)
    eng.langevin_steps(20, MODE_GRAPH)   # the batch keeps running after a swap


@pytest.mark.gpu
def test_energy_of_a_set_equals_the_batch_energies_gpu():
    """tis_energy_sumup_set (the replicas' energies of an exchange) gives
    the bits of the whole-batch energy_sumup for the systems it names, in
    its order and with repeats."""
    from tis_amd.engine import Engine, MODE_GRAPH
    systems = []
    for t in (300.0, 330.0, 360.0):
        s = T.build_rna_system([T.config0_chain()])
        T.set_temperature(s, t, 0.150)
        systems.append(s)
    eng = Engine(systems, keys=[0, 1, 2], seed=3)
    rng_ = np.random.default_rng(4)
    for k, s in enumerate(systems):
        eng.set_coords(k, T.helix_coords(s, rng_))
    eng.langevin_init(tstep=0.05, rebuild_every=10)
    eng.langevin_steps(30, MODE_GRAPH)
    eng.force_all()
    eall = eng.energy_all()
    es = eng.energy_systems([2, 0, 2, 1])
    assert np.array_equal(es, eall[[2, 0, 2, 1]])
    assert np.array_equal(eng.energy_systems([1]), eall[[1]])


@pytest.mark.gpu
def test_label_apply_equals_set_temperature_gpu():
    """A device label swap gives the bits of a host parameter refresh
    (tis_set_system_T), velocities aside."""
    from tis_amd.engine import Engine, MODE_GRAPH
    temps = [300.0, 340.0]
    systems = []
    for t in temps:
        s = T.build_rna_system([T.config0_chain()])
        T.set_temperature(s, t, 0.150)
        systems.append(s)
    x = T.helix_coords(systems[0], np.random.default_rng(4))
    out = []
    for route in ("host", "label"):
        eng = Engine([systems[0], systems[0]], keys=[0, 1], seed=3)
        for k in range(2):
            eng.set_coords(k, x)
        eng.langevin_init(tstep=0.05, rebuild_every=10)
        eng.langevin_steps(40, MODE_GRAPH)
        if route == "host":
            eng.set_temperature_system(1, systems[1])
        else:
            eng.label_define(0, 0, systems[0])
            eng.label_define(1, 0, systems[1])
            eng.label_apply([1], [1])
        eng.langevin_steps(60, MODE_GRAPH)
        out.append([eng.coords(1), eng.get_state(1, 1), eng.energy_sumup(1)])
        eng.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
