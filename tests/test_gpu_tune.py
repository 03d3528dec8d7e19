"""GPU tests of the force pass's launch shapes and of tis_tune: every shape
gives the same forces and trajectories bit for bit, and the tuner changes
neither the state nor the counters."""
import itertools

import numpy as np
import pytest

from tis_amd.engine import Engine, MODE_GRAPH

from systems import compact_allpairs, hairpin_design, tstack_hairpin, valence_crowd

pytestmark = pytest.mark.gpu

SHAPES = list(itertools.product((1, 2), (1, 2), (1, 0)))


def _batch():
    return [hairpin_design(3, 300), hairpin_design(4, 90), compact_allpairs(seed=12, ions=True),
            tstack_hairpin(), valence_crowd(nc=20)]


def _run(shape):
    cases = _batch()
    eng = Engine([s for s, _ in cases], keys=list(range(len(cases))), seed=17)
    for k, (_, x) in enumerate(cases):
        eng.set_coords(k, x)
    if shape is not None:
        assert eng.launch_shape(shape) == tuple(shape)
    eng.neighbor(-1)
    f = eng.force_all()
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(20, MODE_GRAPH)
    st = [[eng.get_state(k, w) for w in range(3)] for k in range(len(cases))]
    cnt = eng.counters()
    eng.close()
    return f, st, cnt


def test_every_launch_shape_is_bitwise_the_same():
    f0, s0, c0 = _run(None)
    for shape in SHAPES:
        print(f"[shape] {shape}", flush=True)
        f, st, cnt = _run(shape)
        assert all(np.array_equal(a, b) for a, b in zip(f, f0)), shape
        for a, b in zip(st, s0):
            assert all(np.array_equal(p, q) for p, q in zip(a, b)), shape
        assert cnt == c0, shape


def test_tune_keeps_state_and_counters():
    cases = _batch()
    eng = Engine([s for s, _ in cases], keys=list(range(len(cases))), seed=17)
    for k, (_, x) in enumerate(cases):
        eng.set_coords(k, x)
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(10, MODE_GRAPH)
    before = [[eng.get_state(k, w) for w in range(3)] for k in range(len(cases))]
    c_before = eng.counters()
    res = eng.tune(reps=4)
    print(f"[tune] {res}", flush=True)
    assert eng.launch_shape() == (res["pairs_per_wave"], res["terms_per_lane"], int(res["offdiag_first"]))
    ms = list(res["candidates_ms"].values())
    assert len(ms) == 8 and min(ms) > 0.0
    assert eng.counters() == c_before
    after = [[eng.get_state(k, w) for w in range(3)] for k in range(len(cases))]
    for a, b in zip(before, after):
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
    with pytest.raises(Exception):
        eng.launch_shape((3, 1, 1))
    eng.close()
