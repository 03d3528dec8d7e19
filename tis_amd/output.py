"""Restart files, energy time series and DCD trajectories of DT15 systems.

Formats follow the reference's writers so that its tools read them:

* restart (.rst): the stream-access binary blocks of src/write_rst.F90:64-110
  (a block is int32 id, int32 payload bytes, payload): STEP (int32
  istep_sim, int64 istep), XYZ / VELO / ACCEL (int32 replica, int32 nmp,
  nmp x 3 FP64 in bead order, x fastest), DTRNA15 (int32 replica, int32
  nhb, int32 nst, then nhb + nst 4-byte logicals of hb_status and
  st_status) and REPLICA (int32 n, then (replica, label) pairs).
  read_rst skips blocks it does not know by their size, as
  src/read_rst.F90 does (sub_skip_block).
* time series (.ts): the columns of src/write_tseries.F90 for the DT15
  model (step, tempk, [label], radg, etot, velet, qscore, rmsd, local, go,
  repul, stack, tstack, hbond, thbond, elect) with the Fortran edit
  descriptors of src/format.F90:32-76; one line for the system and, with
  several units, one per unit (i_output_energy_style 0) from energy_unit.
  The radius of gyration restates src/simu_radiusg.F90 (unweighted, per
  unit about its own centre, the whole system about the common centre and
  0 when an ion unit is present).
* DCD: the CHARMM layout of src/write_xyz_dcd.F90 (84-byte header block,
  title block, atom count, then per frame X, Y, Z as FP32 records).

checkpoint / restore move one system of an Engine through a restart file;
a run split at a list-rebuild boundary and resumed from the file continues
bit for bit (the counter-based noise depends only on the step number).
"""
from __future__ import annotations

import struct

import numpy as np

from .sysdesc import (CLASS_ION, E_BANGLE, E_BOND, E_ELE, E_EXV, E_HBOND, E_STACK, E_THBOND, E_TOTAL,
                      E_TSTACK, E_VELO)

RST_STEP, RST_XYZ, RST_VELO, RST_ACCEL, RST_REPLICA, RST_DTRNA15 = 1, 2, 3, 4, 5, 6
HIGH_ENERGY_JUDGE = 1.0e9        # const_physical.F90
HIGH_ENERGY_OUT = 99999.99       # write_tseries.F90


# ----------------------------------------------------------------- restart
def _block(bid: int, payload: bytes) -> bytes:
    return struct.pack("<ii", bid, len(payload)) + payload


def write_rst(path, xyz, velo, accel, istep: int, istep_sim: int = 1, replica: int = 1,
              hb_status=None, st_status=None, rep2lab=None):
    """One replica's restart file (blocks of write_rst.F90; replica is the
    1-based grep)."""
    x = np.ascontiguousarray(xyz, dtype="<f8").reshape(-1, 3)
    n = len(x)
    out = []
    if rep2lab is not None:
        pay = struct.pack("<i", len(rep2lab)) + b"".join(struct.pack("<ii", i + 1, int(l) + 1)
                                                         for i, l in enumerate(rep2lab))
        out.append(_block(RST_REPLICA, pay))
    out.append(_block(RST_STEP, struct.pack("<iq", int(istep_sim), int(istep))))
    for bid, arr in ((RST_XYZ, x), (RST_VELO, velo), (RST_ACCEL, accel)):
        a = np.ascontiguousarray(arr, dtype="<f8").reshape(n, 3)
        out.append(_block(bid, struct.pack("<ii", int(replica), n) + a.tobytes()))
    if hb_status is not None or st_status is not None:
        hb = np.asarray(hb_status if hb_status is not None else [], dtype=bool)
        st = np.asarray(st_status if st_status is not None else [], dtype=bool)
        pay = struct.pack("<iii", int(replica), len(hb), len(st))
        pay += hb.astype("<i4").tobytes() + st.astype("<i4").tobytes()
        out.append(_block(RST_DTRNA15, pay))
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def read_rst(path) -> dict:
    """Blocks of a restart file: istep_sim, istep, xyz, velo, accel (N, 3),
    replica, hb_status, st_status (bool arrays) and rep2lab (0-based) where
    present; other blocks are skipped by their size."""
    data = open(path, "rb").read()
    out, p = {}, 0
    while p + 8 <= len(data):
        bid, size = struct.unpack_from("<ii", data, p)
        p += 8
        pay = data[p:p + size]
        if len(pay) != size:
            raise ValueError(f"{path}: block {bid} truncated")
        p += size
        if bid == RST_STEP:
            out["istep_sim"], out["istep"] = struct.unpack_from("<iq", pay)
        elif bid in (RST_XYZ, RST_VELO, RST_ACCEL):
            rep, n = struct.unpack_from("<ii", pay)
            arr = np.frombuffer(pay, dtype="<f8", count=3 * n, offset=8).reshape(n, 3).copy()
            out[{RST_XYZ: "xyz", RST_VELO: "velo", RST_ACCEL: "accel"}[bid]] = arr
            out["replica"] = rep
        elif bid == RST_DTRNA15:
            rep, nhb, nst = struct.unpack_from("<iii", pay)
            v = np.frombuffer(pay, dtype="<i4", count=nhb + nst, offset=12)
            out["hb_status"], out["st_status"] = v[:nhb] != 0, v[nhb:] != 0
        elif bid == RST_REPLICA:
            (nrep,) = struct.unpack_from("<i", pay)
            pairs = np.frombuffer(pay, dtype="<i4", count=2 * nrep, offset=4).reshape(nrep, 2)
            out["rep2lab"] = [int(l) - 1 for _, l in pairs]
    return out


def checkpoint(eng, isys: int, path, istep: int):
    """Write system isys of an Engine (coordinates, velocities,
    accelerations, hbond and stack states) at its step istep."""
    hb, st = eng.term_status(isys)
    write_rst(path, eng.get_state(isys, 0), eng.get_state(isys, 1), eng.get_state(isys, 2), istep,
              hb_status=hb, st_status=st)


def restore(eng, isys: int, path) -> int:
    """Load a restart file into system isys; returns its istep.  The
    engine's next tis_langevin_init must keep the loaded accelerations
    (init_accel=False) and start at that step for the run to continue."""
    r = read_rst(path)
    n = eng.systems[isys].nmp
    for which, key in ((0, "xyz"), (1, "velo"), (2, "accel")):
        if r[key].shape != (n, 3):
            raise ValueError(f"{path}: {key} has {len(r[key])} beads, system {isys} has {n}")
        eng.set_state(isys, which, r[key])
    return int(r["istep"])


# ----------------------------------------------------------------- radius of gyration
def radius_of_gyration(system, xyz):
    """(rg of the whole system, rg per unit) as simu_radiusg.F90 computes them."""
    x = np.asarray(xyz, dtype=np.float64)
    units = np.asarray(system.unit)
    nu = system.nunit
    rg_unit = np.zeros(nu)
    ion = [bool(np.all(np.asarray(system.iclass)[units == u] == CLASS_ION)) for u in range(nu)]
    keep = np.zeros(len(x), bool)
    for u in range(nu):
        if ion[u]:
            continue
        m = units == u
        keep |= m
        c = x[m].mean(axis=0)
        rg_unit[u] = float(np.sqrt(((x[m] - c) ** 2).sum() / m.sum()))
    if any(ion):
        return 0.0, rg_unit
    c = x[keep].sum(axis=0) / len(x)
    return float(np.sqrt(((x[keep] - c) ** 2).sum() / len(x))), rg_unit


# ----------------------------------------------------------------- time series
def _clip(v):
    return HIGH_ENERGY_OUT if v > HIGH_ENERGY_JUDGE else v


class TimeSeries:
    """A .ts file of one system (write_tseries.F90, DT15 columns,
    i_output_energy_style 0)."""

    def __init__(self, path, system, with_ele=True, replica=False):
        self.fh = open(path, "w")
        self.sys = system
        self.ele = with_ele
        self.replica = replica
        self.header_done = False

    def _cols(self):
        cols = [("step", "%13s"), ("tempk", "%9s")]
        if self.replica:
            cols.append(("label", "%6s"))
        cols += [("radg", "%9s"), ("etot", "%11s"), ("velet", "%11s"), ("qscore", "%7s"), ("rmsd", "%9s"),
                 ("local", "%11s"), ("go", "%11s"), ("repul", "%11s"), ("stack", "%11s"), ("tstack", "%11s"),
                 ("hbond", "%11s"), ("thbond", "%11s")]
        if self.ele:
            cols.append(("elect", "%11s"))
        return cols

    def header(self, energy):
        w = self.fh.write
        w("# initial_energy\n# total_energy = %11.4f\n# t_series\n\n" % energy[E_TOTAL])
        w("#########################################################\n")
        names = "".join(f % n for n, f in self._cols())
        w("#    " + names + "\n")
        w("#unit" + names + "\n")
        w("#########################################################\n\n")
        self.header_done = True

    def _line(self, head, step, tempk, label, rg, e):
        parts = [head.ljust(5)[:5], "%13d" % step, "%9.2f" % tempk]
        if self.replica:
            parts.append("%6d" % label)
        local = e[E_BOND] + e[E_BANGLE]
        parts += ["%9.2f" % _clip(rg), "%11.2f" % _clip(e[E_TOTAL]), "%11.2f" % _clip(e[E_VELO]),
                  "%7.3f" % 0.0, "%9.2f" % 0.0, "%11.2f" % _clip(local), "%11.2f" % 0.0,
                  "%11.2f" % _clip(e[E_EXV]), "%11.2f" % _clip(e[E_STACK]), "%11.2f" % e[E_TSTACK],
                  "%11.2f" % _clip(e[E_HBOND]), "%11.2f" % e[E_THBOND]]
        if self.ele:
            parts.append("%11.2f" % _clip(e[E_ELE]))
        return "".join(parts) + "\n"

    def write(self, step, tempk, energy, xyz, energy_unit=None, label=0):
        """One record: energy[25] of the system and, with several units,
        energy_unit (nunit, nunit, 25) after style-0 post-processing."""
        if not self.header_done:
            self.header(energy)
        rg, rg_unit = radius_of_gyration(self.sys, xyz)
        self.fh.write(self._line("", step, tempk, label, rg, energy))
        if self.sys.nunit > 1 and energy_unit is not None:
            for u in range(self.sys.nunit):
                self.fh.write(self._line("#%d" % (u + 1), step, tempk, label, rg_unit[u], energy_unit[u, u]))
        self.fh.flush()

    def close(self):
        self.fh.close()


# ----------------------------------------------------------------- DCD
class DCDWriter:
    """CHARMM DCD trajectory as write_xyz_dcd.F90 writes it (coordinates
    "CORD" or velocities "VELD")."""

    def __init__(self, path, system, nset, istart, nsave, nstep, tstep, tempk, velocities=False):
        self.fh = open(path, "wb")
        self.n = system.nmp
        units = np.asarray(system.unit)
        last = [int(np.nonzero(units == u)[0].max()) + 1 for u in range(system.nunit)]
        w = self.fh.write
        hdr = b"VELD" if velocities else b"CORD"
        body = hdr + struct.pack("<9i", nset, istart, nsave, nstep, system.nunit, 0, 0, 0, 0)
        body += struct.pack("<f", tstep) + struct.pack("<9i", *([0] * 9)) + struct.pack("<i", 24)
        w(struct.pack("<i", 84) + body + struct.pack("<i", 84))
        titles = ["=" * 25 + " Molecular Dynamics " + "=" * 35,
                  "==== MI355X DT15 engine; format of CafeMol 2.1 DCD output ====",
                  "%r" % float(tempk)] + ["%6d" % v for v in last]
        tb = b"".join(t.encode()[:80].ljust(80) for t in titles)
        nb = 4 + 80 * len(titles)
        w(struct.pack("<ii", nb, len(titles)) + tb + struct.pack("<i", nb))
        w(struct.pack("<iii", 4, self.n, 4))

    def frame(self, xyz):
        a = np.asarray(xyz, dtype=np.float64).reshape(self.n, 3)
        nb = struct.pack("<i", 4 * self.n)
        for d in range(3):
            self.fh.write(nb + a[:, d].astype("<f4").tobytes() + nb)

    def close(self):
        self.fh.close()


def read_dcd(path):
    """(header dict, frames (nframe, N, 3) FP32) of a DCD file written as above."""
    data = open(path, "rb").read()
    (b0,) = struct.unpack_from("<i", data, 0)
    hdr = data[4:4 + b0]
    nset, istart, nsave, nstep, nunit = struct.unpack_from("<5i", hdr, 4)
    (tstep,) = struct.unpack_from("<f", hdr, 40)
    p = 8 + b0
    (tb,) = struct.unpack_from("<i", data, p)
    (ntitle,) = struct.unpack_from("<i", data, p + 4)
    titles = [data[p + 8 + 80 * k:p + 88 + 80 * k].decode().rstrip() for k in range(ntitle)]
    p += 8 + tb
    (natom,) = struct.unpack_from("<i", data, p + 4)
    p += 12
    frames = []
    while p < len(data):
        fr = np.zeros((natom, 3), np.float32)
        for d in range(3):
            fr[:, d] = np.frombuffer(data, dtype="<f4", count=natom, offset=p + 4)
            p += 8 + 4 * natom
        frames.append(fr)
    return ({"hdr": hdr[:4].decode(), "nset": nset, "istart": istart, "nsave": nsave, "nstep": nstep,
             "nunit": nunit, "tstep": tstep, "titles": titles, "natom": natom},
            np.asarray(frames).reshape(-1, natom, 3))
