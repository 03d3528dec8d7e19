"""GPU tests of checkpoint / restore through restart files and of the hbond
and stack states recorded in them."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tis_amd import output as O
from tis_amd.engine import Engine, MODE_EAGER, MODE_GRAPH

from systems import compact_allpairs, hairpin_design, tstack_hairpin

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_EAGER])
def test_checkpoint_resume_is_bitwise(tmp_path, mode):
    """A 60-step run equals, bit for bit, 30 steps + a restart file + a new
    engine restored from it + 30 steps (the split at a list rebuild)."""
    systems = [hairpin_design(500, 120), compact_allpairs(seed=43)]
    path = tmp_path / "c.rst"
    for k, (s, x) in enumerate(systems):
        print(f"[checkpoint] system {k} N={s.nmp} mode={mode}", flush=True)
        full = Engine([s], keys=[3 + k], seed=9)
        full.set_coords(0, x)
        full.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
        full.langevin_steps(60, mode)
        xf, vf = full.coords(0), full.get_state(0, 1)
        full.close()
        part = Engine([s], keys=[3 + k], seed=9)
        part.set_coords(0, x)
        part.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
        part.langevin_steps(30, mode)
        O.checkpoint(part, 0, path, istep=30)
        part.close()
        res = Engine([s], keys=[3 + k], seed=9)
        istep = O.restore(res, 0, path)
        assert istep == 30
        res.langevin_init(tstep=0.05, rebuild_every=10, step0=istep, init_accel=False)
        res.langevin_steps(30, mode)
        assert np.array_equal(res.coords(0), xf)
        assert np.array_equal(res.get_state(0, 1), vf)
        res.close()


def test_term_status_matches_oracle():
    """hb_status of the last force evaluation (after the valence pass) and
    st_status of the tertiary-stack switch-off, against the oracle."""
    for s, x in (tstack_hairpin(), compact_allpairs(seed=12)):
        eng = Engine([s], keys=[0], seed=0)
        eng.set_coords(0, x)
        eng.neighbor(0)
        eng.force_sumup(0)
        hb, st = eng.term_status(0)
        eng.close()
        o = oracle.Oracle(s)
        o.neighbor(x)
        _, shb, _ = o.term("hbond", x)
        o.force(x, seed=0, sys_id=0, step=0)
        ost = np.ones(max(len(s.st_mp), 1), np.uint8)
        f = np.zeros_like(x)
        oracle.lib().oracle_force_tstack(C.byref(o.cs), oracle.fp(x), oracle.fp(x),
                                         ost.ctypes.data_as(C.c_char_p), oracle.fp(f))
        assert np.array_equal(st, ost[:len(s.st_mp)] != 0)
        # the oracle's hbond term with the force pass's draws (seed 0, step 0)
        ws_status = np.zeros(len(s.hb_mp), bool)
        nh = max(len(s.hb_mp), 1)
        forh, ene, stat = np.zeros(18 * nh), np.zeros(nh), np.zeros(nh, np.uint8)
        hbl = np.ascontiguousarray(o.lists(2), dtype=np.int32)
        L = oracle.lib()
        L.oracle_hbond_eval(C.byref(o.cs), oracle.fp(x), oracle.fp(x), oracle.ip(hbl), len(hbl),
                            oracle.fp(forh), oracle.fp(ene), stat.ctypes.data_as(C.c_char_p))
        L.oracle_hbond_valence(C.byref(o.cs), oracle.ip(hbl), len(hbl), oracle.fp(ene),
                               stat.ctypes.data_as(C.c_char_p), 0, 0, 0)
        ws_status[:] = stat[:len(s.hb_mp)] != 0
        assert np.array_equal(hb, ws_status)
        del shb
