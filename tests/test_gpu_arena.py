"""GPU tests of the batch arena (tis_system_append / tis_system_retire):
systems appended to a running batch, retired ones, and a new layout that
carries the live systems' state on the device all leave every live system
bitwise where it would be alone."""
import numpy as np
import pytest

from tis_amd.engine import Engine, MODE_EAGER, MODE_GRAPH

from systems import bwyv19, compact_allpairs, hairpin_design, small_ions_box, tstack_hairpin

pytestmark = pytest.mark.gpu

SEED = 4242


def _alone(s, x, key, steps, mode=MODE_GRAPH):
    e = Engine([s], keys=[key], seed=SEED)
    e.set_coords(0, x)
    e.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    e.langevin_steps(steps, mode)
    out = [e.get_state(0, w) for w in range(3)]
    e.close()
    return out


def _admit(eng, idx, x, origin):
    eng.set_state(idx, 0, x)
    eng.set_state(idx, 1, np.zeros_like(x))
    eng.set_state(idx, 2, np.zeros_like(x))
    eng.system_origin(idx, origin, True)


@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_EAGER])
def test_append_retire_relayout_bitwise(mode):
    """Start with two systems, run 20 steps; append two (a new layout: the
    batch was sized exactly), run 20; retire one, append one (in place),
    run 20; append a large one (a second layout, dropping the retired
    storage), run 20.  Every live system's x, v, a equal its run alone over
    the same number of its own steps, bit for bit."""
    cases = [hairpin_design(3, 120), tstack_hairpin(), compact_allpairs(seed=8, ions=True),
             hairpin_design(9, 200), hairpin_design(13, 100), bwyv19(), hairpin_design(11, 600)]
    keys = [10 + k for k in range(len(cases))]
    plan = [([0, 1], []), ([2, 3], []), ([4], [1]), ([5, 6], [])]   # (append, retire) per phase
    steps = 20
    eng = None
    idx, origin = {}, {}
    layouts = []
    g = 0
    for phase, (add, drop) in enumerate(plan):
        print(f"[arena] phase {phase}: append {add}, retire {drop}, mode {mode}", flush=True)
        if drop:
            eng.retire([idx[k] for k in drop])
            for k in drop:
                idx.pop(k)
        if eng is None:
            eng = Engine([cases[k][0] for k in add], keys=[keys[k] for k in add], seed=SEED)
            new = list(range(len(add)))
        else:
            new = eng.append([cases[k][0] for k in add], keys=[keys[k] for k in add])
        for k, i in zip(add, new):
            idx[k] = i
            origin[k] = g
            _admit(eng, i, cases[k][1], g)
        for k in idx:
            if origin[k] < g:
                eng.system_origin(idx[k], origin[k], False)
        eng.langevin_init(tstep=0.05, rebuild_every=10, step0=g, init_accel=None)
        eng.langevin_steps(steps, mode)
        g += steps
        print(f"   {eng.arena_stats()}", flush=True)
        layouts.append(eng.arena_stats()["layouts"])
    st = eng.arena_stats()
    # phase 1 and 3 lay the arena out again (the retired storage goes), the
    # phase-2 system fits in place
    assert layouts == [1, 2, 2, 3], layouts
    assert st["live_systems"] == 6 and st["indices"] == 7
    assert st["stored_beads"] == st["live_beads"]
    tb = eng.transfer_bytes()
    assert tb["d2d_state"] > 0
    for k, i in idx.items():
        got = [eng.get_state(i, w) for w in range(3)]
        want = _alone(cases[k][0], cases[k][1], keys[k], g - origin[k], mode)
        for a, b, what in zip(got, want, "xva"):
            assert np.array_equal(a, b), f"system {k} {what} differs"
    with pytest.raises(Exception):
        eng.get_state(1, 0)     # retired
    eng.close()


def test_append_periodic_forces_a_layout():
    """A periodic system appended to a non-periodic batch needs the wrapped
    position arrays: the batch is laid out again, and forces of every
    system equal those of the system alone."""
    a, b = hairpin_design(5, 80), small_ions_box()
    eng = Engine([a[0]], keys=[1], seed=SEED)
    eng.set_coords(0, a[1])
    (i,) = eng.append([b[0]], keys=[2])
    eng.set_coords(i, b[1])
    assert eng.arena_stats()["layouts"] == 2
    eng.neighbor(-1)
    f = eng.force_all()
    for k, (s, x) in enumerate((a, b)):
        e = Engine([s], keys=[k + 1], seed=SEED)
        e.set_coords(0, x)
        e.neighbor(0)
        assert np.array_equal(e.force_sumup(0), f[k])
        e.close()
    eng.close()
