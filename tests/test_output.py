"""CPU tests of the restart, time-series and DCD writers: round trips, and
the restart file read back by the reference's own show_rst program
(oracle/_ref/show_rst, compiled from src/show_rst.F90) where it is built."""
import os
import subprocess

import numpy as np
import pytest

from tis_amd import output as O
from tis_amd.sysdesc import E_BOND, E_EXV, E_TOTAL, E_TSTACK

from systems import bwyv19, two_chain_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHOW_RST = os.path.join(ROOT, "oracle", "_ref", "show_rst")


def _state(n=7, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(3, n, 3)) * np.array([30.0, 1.0, 0.1])[:, None, None]


def test_rst_round_trip(tmp_path):
    x, v, a = _state()
    p = tmp_path / "s.rst"
    O.write_rst(p, x, v, a, istep=2 ** 40 + 3, istep_sim=2, replica=3, hb_status=[1, 0, 1, 1],
                st_status=[0, 1], rep2lab=[2, 0, 1])
    r = O.read_rst(p)
    assert r["istep"] == 2 ** 40 + 3 and r["istep_sim"] == 2 and r["replica"] == 3
    assert np.array_equal(r["xyz"], x) and np.array_equal(r["velo"], v) and np.array_equal(r["accel"], a)
    assert r["hb_status"].tolist() == [True, False, True, True] and r["st_status"].tolist() == [False, True]
    assert r["rep2lab"] == [2, 0, 1]


def test_rst_reader_skips_unknown_blocks(tmp_path):
    """read_rst.F90's sub_skip_block: a block it does not want is skipped by
    its size (here a WIDOM-like block with id 8 ahead of the others)."""
    x, v, a = _state(4, 1)
    p = tmp_path / "s.rst"
    O.write_rst(p, x, v, a, istep=9)
    raw = open(p, "rb").read()
    extra = np.array([8, 12], "<i4").tobytes() + np.array([1], "<i4").tobytes() + np.array([7], "<i8").tobytes()
    open(p, "wb").write(extra + raw)
    r = O.read_rst(p)
    assert r["istep"] == 9 and np.array_equal(r["xyz"], x)


@pytest.mark.skipif(not os.path.exists(SHOW_RST), reason="oracle/_ref/show_rst not built (reference absent)")
def test_rst_read_by_reference_show_rst(tmp_path):
    """The reference's show_rst (src/show_rst.F90, RSTBLK ids of
    const_index.F90:470-480) reads every block of the file back."""
    x, v, a = _state(5, 2)
    p = tmp_path / "s.rst"
    O.write_rst(p, x, v, a, istep=123456789012, hb_status=[1, 0, 1], st_status=[0, 1], rep2lab=[1, 0])
    out = subprocess.run([SHOW_RST, str(p)], capture_output=True, text=True, check=True).stdout
    lines = [ln.strip() for ln in out.splitlines() if ln.strip()]
    assert "istep: 123456789012" in lines and "istep_sim: 1" in lines
    assert "rep2lab: 1 2" in lines and "rep2lab: 2 1" in lines

    def block(name, n):
        k = lines.index(name)
        return np.array([[float(t.replace("D", "E")) for t in ln.split()] for ln in lines[k + 3:k + 3 + n]])

    assert np.array_equal(block("# xyz_mp_rep", 5), x)
    assert np.array_equal(block("# velo_mp", 5), v)
    assert np.array_equal(block("# accel_mp", 5), a)
    assert [ln for ln in lines if ln.startswith("hb_status")] == \
        ["hb_status( 1 ): T", "hb_status( 2 ): F", "hb_status( 3 ): T"]
    assert [ln for ln in lines if ln.startswith("st_status")] == ["st_status( 1 ): F", "st_status( 2 ): T"]


def test_radius_of_gyration():
    """simu_radiusg.F90: unweighted, each unit about its own centre; the
    whole system 0 when an ion unit is present."""
    s, x = two_chain_system()
    rg, rgu = O.radius_of_gyration(s, x)
    assert rg == 0.0 and rgu[-1] == 0.0
    m = np.asarray(s.unit) == 0
    c = x[m].mean(axis=0)
    assert abs(rgu[0] - np.sqrt(((x[m] - c) ** 2).sum(axis=1).mean())) <= 1e-12 * rgu[0]


def test_time_series_columns(tmp_path):
    """The DT15 columns of write_tseries.F90 with the widths of format.F90
    (step 1x,i12; tempk/radg 1x,f8.2; energies 1x,f10.2; qscore 1x,f6.3),
    one line for the system and one per unit."""
    s, x = bwyv19()
    e = np.zeros(25)
    e[E_TOTAL], e[E_BOND], e[E_EXV], e[E_TSTACK] = -123.456, 10.0, 2e10, -4.25
    eu = np.zeros((s.nunit, s.nunit, 25))
    eu[0, 0, E_TOTAL] = -100.0
    p = tmp_path / "s.ts"
    ts = O.TimeSeries(p, s)
    ts.write(1000, 310.15, e, x, eu)
    ts.close()
    lines = open(p).read().splitlines()
    data = [ln for ln in lines if not ln.startswith("#") or ln.startswith("#1") or ln.startswith("#2")]
    data = [ln for ln in data if ln.strip()]
    assert len(data) == 3
    head = [ln for ln in lines if ln.startswith("#    ")][0]
    assert head.split()[1:6] == ["step", "tempk", "radg", "etot", "velet"]
    row = data[0]
    assert row[5:18] == "%13d" % 1000 and row[18:27] == "%9.2f" % 310.15
    assert row[36:47] == "%11.2f" % -123.456
    cols = row[5:].split()
    names = head.split()[1:]
    assert float(cols[names.index("repul")]) == 99999.99   # HIGH_ENERGY_OUT above 1e9
    assert float(cols[names.index("tstack")]) == -4.25
    assert data[1].startswith("#1") and float(data[1][5:].split()[names.index("etot")]) == -100.0


def test_dcd_round_trip(tmp_path):
    s, x = bwyv19()
    p = tmp_path / "t.dcd"
    w = O.DCDWriter(p, s, nset=3, istart=0, nsave=100, nstep=300, tstep=0.05, tempk=310.15)
    frames = [x + k for k in range(3)]
    for f in frames:
        w.frame(f)
    w.close()
    h, fr = O.read_dcd(p)
    assert h["hdr"] == "CORD" and h["natom"] == s.nmp and h["nset"] == 3 and h["nunit"] == s.nunit
    assert abs(h["tstep"] - 0.05) < 1e-7 and h["titles"][3].strip() == "85"
    assert np.array_equal(fr, np.asarray(frames, np.float32))
