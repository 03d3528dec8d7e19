"""The Fortran host path: tis_hip_check (module tis_hip over iso_c_binding)
calls neighbor / force_sumup / energy_sumup / Langevin through the C-ABI;
its results equal the Python host's bit for bit and the oracle's within the
parity tolerances."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tis_amd.build import FORTRAN_EXE
from tis_amd.engine import Engine, MODE_GRAPH

from fortran_io import read_result, write_system
from systems import compact_allpairs, force_error, hairpin64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["hairpin64", "ions_periodic"])
def test_fortran_host_matches(case, tmp_path):
    if not os.path.exists(FORTRAN_EXE):
        pytest.skip("Fortran check program not built (flang absent at build time)")
    s, x = hairpin64(0) if case == "hairpin64" else compact_allpairs(seed=13, ions=True, periodic=True)
    nsteps = 30
    fin, fout = tmp_path / "sys.bin", tmp_path / "out.bin"
    write_system(fin, s, x, nsteps=nsteps, tstep=0.05)
    print(f"[fortran] {case} N={s.nmp}", flush=True)
    res = subprocess.run([FORTRAN_EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    f, e, xs = read_result(fout, s.nmp, nsteps)
    # the Python host over the same library
    eng = Engine([s], seed=0)
    eng.set_coords(0, x)
    eng.neighbor(0)
    fp = eng.force_sumup(0)
    ep = eng.energy_sumup(0)
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(nsteps, MODE_GRAPH)
    assert np.array_equal(f, fp) and np.array_equal(e, ep) and np.array_equal(xs, eng.coords(0))
    o = oracle.Oracle(s)
    o.neighbor(x)
    fo, _ = o.force(x, seed=0, sys_id=0, step=0)
    assert force_error(f, fo).max() <= 1e-5
