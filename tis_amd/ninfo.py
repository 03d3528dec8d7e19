"""Native-info (.ninfo) and coarse-grained PDB input/output for DT15 systems.

The reference takes DT15 topology only from .ninfo files
(src/setp_nativestruct.F90:38-40) read by src/read_nativeinfo.F90:121-1120,
and bead names and coordinates from coarse-grained PDB files
(src/read_pdb_cg.F90).  This module reads and writes the records of the
current format that the DT15 path uses, in the all-in-one layout (global
1-based bead numbers in the imp columns):

  bond      ibd iunit1 iunit2 imp1 imp2 imp1un imp2un bd_nat factor correct coef [type]
  angl      iba iunit1 iunit2 imp1 imp2 imp3 imp1un imp2un imp3un ba_nat(deg) factor correct coef [type]
  bs-dist   ist iunit1 iunit2 imp1 imp2 imp1un imp2un h s Tm dist coef [type]
  bs-dihd   ist idih iunit1 iunit2 imp1..imp4 imp1un..imp4un dih(deg) coef PSPS|SPSP
  hb-dist   ihb iunit1 iunit2 imp1 imp2 imp1un imp2un U0 dist coef S|T nHB a11 a12 [a21 a22 [a31 a32]]
  hb-angl   ihb iba iunit1 iunit2 imp1 imp2 imp3 imp1un imp2un imp3un angle(deg) coef
  hb-dihd   ihb idih iunit1 iunit2 imp1..imp4 imp1un..imp4un dih(deg) coef
  tbs-dist  ist iunit1 iunit2 imp1 imp2 imp1un imp2un U0 dist coef type1 type2
  tbs-angl  ist iba iunit1 iunit2 imp1 imp2 imp3 imp1un imp2un imp3un angle(deg) coef
  tbs-dihd  ist idih iunit1 iunit2 imp1..imp4 imp1un..imp4un dih(deg) coef

and turns them into a System with the reader's conventions: bead roles of a
stack (read_nativeinfo.F90:482-488) and tertiary stack (:665-670), the
role of each hb-angl / hb-dihd / tbs-* record by the beads it names
(:833-1070), hbond sites by atom name (:573-589 and site_local :1285-1336),
degrees to radians ((x/180)*pi for bond angles :231, x*pi/180 for the
others :761,843,907,1000,1052), and the tertiary stack to secondary stack
map of src/mloop_dtrna.F90:73-136.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import topology as T
from .sysdesc import CLASS_ION, CLASS_RNA, MP_BASE, MP_PHOS, MP_SUGAR, System

_BASE_OF = {"Ab": "A", "Cb": "C", "Gb": "G", "Ub": "U"}
_ION_OF = {"MG": "Mg", "K": "K", "CL": "Cl", "NA": "Na", "CA": "Ca"}


# --------------------------------------------------------------------------- PDB
@dataclass
class CGBeads:
    """Beads of coarse-grained PDB files: names in the build's convention
    ('P', 'S', 'A'/'C'/'G'/'U' for bases, ion species), their unit (one per
    TER-terminated chain), residue names and (N, 3) coordinates."""
    names: list
    unit: list
    resname: list
    xyz: np.ndarray


def read_pdb_cg(paths) -> CGBeads:
    """ATOM records of one or more coarse-grained PDB files in the fixed
    columns of src/read_pdb_cg.F90 ('(a6, i5, 1x, a4, a1, a3, a2, i4, 4x,
    3f8.3)'); a TER line or a new file starts a new unit."""
    if isinstance(paths, str):
        paths = [paths]
    names, unit, resname, xyz = [], [], [], []
    u = -1
    for path in paths:
        reading = False
        with open(path) as fh:
            for line in fh:
                if line.startswith("TER") or line.startswith(">>"):
                    reading = False
                    continue
                if not line.startswith("ATOM"):
                    continue
                if not reading:
                    u += 1
                    reading = True
                atom = line[12:16].strip()
                res = line[17:20].strip()
                x = float(line[30:38]); y = float(line[38:46]); z = float(line[46:54])
                if atom in ("P", "S"):
                    nm = atom
                elif atom in _BASE_OF:
                    nm = _BASE_OF[atom]
                elif atom.upper() in _ION_OF:
                    nm = _ION_OF[atom.upper()]
                else:
                    raise ValueError(f"unknown coarse-grained atom {atom!r} in {path}")
                names.append(nm); unit.append(u); resname.append(res); xyz.append((x, y, z))
    return CGBeads(names, unit, resname, np.asarray(xyz, dtype=np.float64).reshape(-1, 3))


def _atom_name(nm):
    if nm in ("P", "S"):
        return nm
    if nm in _BASE_OF.values():
        return nm + "b"
    return {v: k for k, v in _ION_OF.items()}[nm]


ION_NAMES = ("Mg", "Ca", "K", "Na", "Cl")


def write_pdb_cg(system: System, xyz, path, chain_ids="ABCDEFGHIJKLMNOPQRSTUVWXYZ"):
    """Coarse-grained PDB of a system in the columns read_pdb_cg reads, one
    TER after each unit.  A residue starts at the first bead of a unit and
    after each base; it is named after its base (a trailing phosphate takes
    the name of the base before it).  Coordinates have three decimals, as
    in the format."""
    names = bead_names(system)
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    lines = []
    res, rname = 0, "N"
    for i, nm in enumerate(names):
        u = int(system.unit[i])
        first = i == 0 or int(system.unit[i - 1]) != u
        if first:
            res = 0
        if nm in ION_NAMES:
            res += 1
            rname = nm
        else:
            if first or names[i - 1] in _BASE_OF.values():
                res += 1
                j = i
                while j < len(names) and int(system.unit[j]) == u and names[j] in ("P", "S"):
                    j += 1
                if j < len(names) and int(system.unit[j]) == u and names[j] in _BASE_OF.values():
                    rname = "R" + names[j]
        an = _atom_name(nm)
        an = an.ljust(4) if nm in ION_NAMES else (" " + an).ljust(4)
        lines.append("ATOM  %5d %s %-3s %1s%4d    %8.3f%8.3f%8.3f\n"
                     % ((i + 1) % 100000, an, rname[:3], chain_ids[u % len(chain_ids)], res % 10000,
                        xyz[i, 0], xyz[i, 1], xyz[i, 2]))
        if i == len(names) - 1 or int(system.unit[i + 1]) != u:
            lines.append("TER\n")
    with open(path, "w") as fh:
        fh.writelines(lines)


def bead_names(system: System):
    """Bead names of a system from its types and masses (bases by mass,
    ions by species)."""
    out = []
    mass_to_base = {round(T.MASS[b], 5): b for b in "ACGU"}
    ion_of_type = {v: k for k, v in T.ION_MPTYPE.items()}
    for i in range(system.nmp):
        t = int(system.mptype[i])
        if t == MP_PHOS:
            out.append("P")
        elif t == MP_SUGAR:
            out.append("S")
        elif t == MP_BASE:
            out.append(mass_to_base[round(float(system.mass[i]), 5)])
        else:
            out.append(ion_of_type[t])
    return out


# --------------------------------------------------------------------------- ninfo
@dataclass
class NinfoRecords:
    """Records of a .ninfo file as read (bead numbers global and 1-based,
    angles in degrees)."""
    bond: list = field(default_factory=list)      # (imp1, imp2, nat, factor, correct, coef, type)
    angl: list = field(default_factory=list)      # (imp1, imp2, imp3, deg, factor, correct, coef, type)
    bs_dist: list = field(default_factory=list)   # (ist, imp1, imp2, h, s, Tm, dist, coef, type)
    bs_dihd: list = field(default_factory=list)   # (ist, idih, imp1..imp4, deg, coef, type)
    hb_dist: list = field(default_factory=list)   # (ihb, imp1, imp2, U0, dist, coef, S|T, atoms)
    hb_angl: list = field(default_factory=list)   # (ihb, iba, imp1, imp2, imp3, deg, coef)
    hb_dihd: list = field(default_factory=list)   # (ihb, idih, imp1..imp4, deg, coef)
    tbs_dist: list = field(default_factory=list)  # (ist, imp1, imp2, U0, dist, coef, type1, type2)
    tbs_angl: list = field(default_factory=list)  # (ist, iba, imp1, imp2, imp3, deg, coef)
    tbs_dihd: list = field(default_factory=list)  # (ist, idih, imp1..imp4, deg, coef)


def read_ninfo(path) -> NinfoRecords:
    """Parse the records listed in the module docstring; other lines
    (headers, comments, records of other models) are skipped."""
    r = NinfoRecords()
    with open(path) as fh:
        for line in fh:
            w = line.split()
            if not w:
                continue
            key = w[0]
            if key == "bond":
                r.bond.append((int(w[4]), int(w[5]), float(w[8]), float(w[9]), float(w[10]), float(w[11]),
                               w[12] if len(w) > 12 else ""))
            elif key == "angl":
                r.angl.append((int(w[4]), int(w[5]), int(w[6]), float(w[10]), float(w[11]), float(w[12]),
                               float(w[13]), w[14] if len(w) > 14 else ""))
            elif key == "bs-dist":
                r.bs_dist.append((int(w[1]), int(w[4]), int(w[5]), float(w[8]), float(w[9]), float(w[10]),
                                  float(w[11]), float(w[12]), w[13] if len(w) > 13 else ""))
            elif key == "bs-dihd":
                r.bs_dihd.append((int(w[1]), int(w[2]), int(w[5]), int(w[6]), int(w[7]), int(w[8]),
                                  float(w[13]), float(w[14]), w[15]))
            elif key == "hb-dist":
                nhb = int(w[12])
                atoms = tuple((w[13 + 2 * q], w[14 + 2 * q]) for q in range(nhb))
                r.hb_dist.append((int(w[1]), int(w[4]), int(w[5]), float(w[8]), float(w[9]), float(w[10]),
                                  w[11], atoms))
            elif key == "hb-angl":
                r.hb_angl.append((int(w[1]), int(w[2]), int(w[5]), int(w[6]), int(w[7]), float(w[11]),
                                  float(w[12])))
            elif key == "hb-dihd":
                r.hb_dihd.append((int(w[1]), int(w[2]), int(w[5]), int(w[6]), int(w[7]), int(w[8]),
                                  float(w[13]), float(w[14])))
            elif key == "tbs-dist":
                r.tbs_dist.append((int(w[1]), int(w[4]), int(w[5]), float(w[8]), float(w[9]), float(w[10]),
                                   int(w[11]), int(w[12])))
            elif key == "tbs-angl":
                r.tbs_angl.append((int(w[1]), int(w[2]), int(w[5]), int(w[6]), int(w[7]), float(w[11]),
                                   float(w[12])))
            elif key == "tbs-dihd":
                r.tbs_dihd.append((int(w[1]), int(w[2]), int(w[5]), int(w[6]), int(w[7]), int(w[8]),
                                   float(w[13]), float(w[14])))
    return r


def _site_local(base, atom):
    """site_local of read_nativeinfo.F90:1285-1336 (1-based within the bead)."""
    a = atom.upper()
    table = {"O2'": 1, "O4'": 2, "O3'": 3, "O5'": 4, "OP1": 1, "OP2": 2}
    if a in table:
        return table[a]
    if a.startswith("N1") or a.startswith("O2"):
        return 1
    if a.startswith("N2"):
        return 2
    if a.startswith("N3"):
        return 3 if base == "G" else 2
    if a.startswith("O4") or a.startswith("N4") or a.startswith("N6"):
        return 3
    if a.startswith("O6"):
        return 4
    if a.startswith("N7"):
        return 4 if base == "A" else 5
    raise ValueError(f"unknown hbond atom {atom!r} on {base}")


def system_from_ninfo(rec: NinfoRecords, beads: CGBeads, **kwargs) -> System:
    """A DT15 System from .ninfo records and the beads of the PDB files they
    refer to.  kwargs go to topology.assemble_system (temperature, salt,
    electrostatics, box...).  The stack U0 is derived from (h, s, Tm) at the
    system temperature (src/simu_set_dtrna.F90:46-56)."""
    names = list(beads.names)
    n = len(names)
    unit = list(beads.unit)
    nunit = max(unit) + 1
    unit_class = [CLASS_ION if all(names[i] in ION_NAMES for i in range(n) if unit[i] == u) else CLASS_RNA
                  for u in range(nunit)]
    tempk = kwargs.get("tempk", 310.0)
    z = lambda k: k - 1  # noqa: E731  1-based -> 0-based
    topo = {}
    topo["bd_mp"] = [(z(a), z(b)) for a, b, *_ in rec.bond]
    topo["bd_nat"] = [x[2] for x in rec.bond]
    topo["bd_coef"] = [x[5] for x in rec.bond]
    topo["ba_mp"] = [(z(a), z(b), z(c)) for a, b, c, *_ in rec.angl]
    topo["ba_nat"] = [(x[3] / 180.0) * T.F_PI for x in rec.angl]
    topo["ba_coef"] = [x[6] for x in rec.angl]
    # consecutive stacks, indexed by ist (read_nativeinfo.F90:438-497, 750-829)
    nst = max((x[0] for x in rec.bs_dist), default=0)
    st_mp = np.zeros((nst, 7), np.int64)
    st_nat = np.zeros((nst, 3))
    st_coef = np.zeros((nst, 4))
    st_hsTm = np.zeros((nst, 3))
    for ist, i1, i2, h, s, tm, dist, coef, _ in rec.bs_dist:
        k = ist - 1
        st_mp[k] = (z(i1), z(i2), z(i1) - 2, z(i1) - 1, z(i1) + 1, z(i1) + 2, z(i1) + 4)
        st_nat[k, 0] = dist
        st_coef[k, 0] = T.stack_u0(h, s, tm, tempk)
        st_coef[k, 1] = coef
        st_hsTm[k] = (h, s, tm)
    for ist, _, a, b, c, d, deg, coef, typ in rec.bs_dihd:
        k = ist - 1
        dih = deg * T.F_PI / 180.0
        if typ == "PSPS":
            if not ((a, b, c, d) == tuple(st_mp[k, 2:6] + 1) or (d, c, b, a) == tuple(st_mp[k, 2:6] + 1)):
                st_mp[k, 2:6] = (z(a), z(b), z(c), z(d))
            st_nat[k, 1] = dih
            st_coef[k, 2] = coef
        elif typ == "SPSP":
            if not ((a, b, c, d) == tuple(st_mp[k, 3:7] + 1) or (d, c, b, a) == tuple(st_mp[k, 3:7] + 1)):
                st_mp[k, 6] = z(d)
            st_nat[k, 2] = dih
            st_coef[k, 3] = coef
        else:
            raise ValueError("unknown type of bs-dihd: " + typ)
    topo.update(st_mp=st_mp, st_nat=st_nat, st_coef=st_coef, st_hsTm=st_hsTm)
    # hydrogen bonds (read_nativeinfo.F90:506-619, 833-986)
    offset, _ = T.hbsite_layout(names)
    nhb = max((x[0] for x in rec.hb_dist), default=0)
    hb_mp = -np.ones((nhb, 6), np.int64)
    hb_nat = np.zeros((nhb, 6))
    hb_coef = np.zeros((nhb, 7))
    hb_site = -np.ones((nhb, 6), np.int64)
    hb_tert = np.zeros(nhb, np.int32)
    for ihb, i1, i2, u0, dist, coef, typ, atoms in rec.hb_dist:
        k = ihb - 1
        a, b = (i1, i2) if i1 < i2 else (i2, i1)
        if i2 < i1:
            atoms = tuple((y, x) for x, y in atoms)
        hb_mp[k, 0], hb_mp[k, 1] = z(a), z(b)
        hb_nat[k, 0] = dist
        hb_coef[k, 0], hb_coef[k, 1] = u0, coef
        for q, (x1, x2) in enumerate(atoms):
            hb_site[k, 2 * q] = offset[z(a)] + _site_local(names[z(a)], x1) - 1
            hb_site[k, 2 * q + 1] = offset[z(b)] + _site_local(names[z(b)], x2) - 1
        if typ not in ("S", "T"):
            raise ValueError("hb-dist type must be S or T")
        hb_tert[k] = 1 if typ == "T" else 0
    for ihb, _, a, b, c, deg, coef in rec.hb_angl:
        k = ihb - 1
        if a > c:
            a, c = c, a
        a, b, c = z(a), z(b), z(c)
        ang = deg * T.F_PI / 180.0
        if b == hb_mp[k, 0] and c == hb_mp[k, 1]:
            _fill(hb_mp, k, 2, a)
            hb_nat[k, 1], hb_coef[k, 2] = ang, coef
        elif a == hb_mp[k, 0] and b == hb_mp[k, 1]:
            _fill(hb_mp, k, 3, c)
            hb_nat[k, 2], hb_coef[k, 3] = ang, coef
        else:
            raise ValueError(f"hb-angl {ihb} does not name its hbond's beads")
    for ihb, _, a, b, c, d, deg, coef in rec.hb_dihd:
        k = ihb - 1
        if a > d:
            a, b, c, d = d, c, b, a
        a, b, c, d = z(a), z(b), z(c), z(d)
        dih = deg * T.F_PI / 180.0
        if b == hb_mp[k, 0] and c == hb_mp[k, 1]:
            _fill(hb_mp, k, 2, a); _fill(hb_mp, k, 3, d)
            hb_nat[k, 3], hb_coef[k, 4] = dih, coef
        elif c == hb_mp[k, 0] and d == hb_mp[k, 1]:
            _fill(hb_mp, k, 2, b)
            hb_mp[k, 4] = a
            hb_nat[k, 4], hb_coef[k, 5] = dih, coef
        elif a == hb_mp[k, 0] and b == hb_mp[k, 1]:
            _fill(hb_mp, k, 3, c)
            hb_mp[k, 5] = d
            hb_nat[k, 5], hb_coef[k, 6] = dih, coef
        else:
            raise ValueError(f"hb-dihd {ihb} does not name its hbond's beads")
    if nhb and (hb_mp < 0).any():
        raise ValueError("an hbond lacks its angle or dihedral records")
    topo.update(hb_mp=hb_mp, hb_nat=hb_nat, hb_coef=hb_coef, hb_site=hb_site, hb_tertiary=hb_tert)
    # tertiary stacks (read_nativeinfo.F90:624-705, 990-1070)
    ntst = max((x[0] for x in rec.tbs_dist), default=0)
    tst_mp = np.zeros((ntst, 6), np.int64)
    tst_nat = np.zeros((ntst, 6))
    tst_coef = np.zeros((ntst, 7))
    side = np.zeros((ntst, 2), np.int64)
    excl = np.zeros((ntst, 2), np.int64)
    for ist, i1, i2, u0, dist, coef, t1, t2 in rec.tbs_dist:
        k = ist - 1
        if i1 > i2:
            i1, i2, t1, t2 = i2, i1, t2, t1
        b1, b2 = z(i1), z(i2)
        tst_mp[k] = (b1, b2, b1 - 1, b2 - 1, b1 + 1, b2 + 1)
        tst_nat[k, 0] = dist
        tst_coef[k, 0], tst_coef[k, 1] = u0, coef
        for q, t in enumerate((t1, t2)):
            if abs(t) not in (1, 2):
                raise ValueError("tbs-dist type must be +-1 or +-2")
            side[k, q] = 1 if t < 0 else 2
            excl[k, q] = 1 if abs(t) == 2 else 0
    for ist, _, a, b, c, deg, coef in rec.tbs_angl:
        k = ist - 1
        if a > c:
            a, c = c, a
        m = tuple(tst_mp[k] + 1)
        ang = deg * T.F_PI / 180.0
        if (a, b, c) == (m[2], m[0], m[1]):
            tst_nat[k, 1], tst_coef[k, 2] = ang, coef
        elif (a, b, c) == (m[0], m[1], m[3]):
            tst_nat[k, 2], tst_coef[k, 3] = ang, coef
        else:
            raise ValueError(f"tbs-angl {ist} does not name its stack's beads")
    for ist, _, a, b, c, d, deg, coef in rec.tbs_dihd:
        k = ist - 1
        if a > d:
            a, b, c, d = d, c, b, a
        m = tuple(tst_mp[k] + 1)
        dih = deg * T.F_PI / 180.0
        if (a, b, c, d) == (m[2], m[0], m[1], m[3]):
            tst_nat[k, 3], tst_coef[k, 4] = dih, coef
        elif (a, b, c, d) == (m[4], m[2], m[0], m[1]):
            tst_nat[k, 4], tst_coef[k, 5] = dih, coef
        elif (a, b, c, d) == (m[0], m[1], m[3], m[5]):
            tst_nat[k, 5], tst_coef[k, 6] = dih, coef
        else:
            raise ValueError(f"tbs-dihd {ist} does not name its stack's beads")
    tst2st, excl = tertiary_stack_map(st_mp, tst_mp, side, excl)
    topo.update(tst_mp=tst_mp, tst_nat=tst_nat, tst_coef=tst_coef, tst2st=tst2st, tst_excl=excl)
    seq = "".join(nm for nm in names if nm in "ACGU" and len(nm) == 1)
    return T.assemble_system(names, unit, unit_class, topo, seq=seq, **kwargs)


def _fill(mp, k, col, val):
    """Set a bead role named by an angle/dihedral record, or check it."""
    if mp[k, col] < 0:
        mp[k, col] = val
    elif mp[k, col] != val:
        raise ValueError(f"record {k + 1}: bead {val + 1} conflicts with {mp[k, col] + 1}")


def tertiary_stack_map(st_mp, tst_mp, side, excl):
    """idtrna_tst2st of src/mloop_dtrna.F90:73-136: for each side of each
    tertiary stack, the consecutive stack of its base with the base three
    beads before (side 1) or after (side 2); when there is none the side is
    made non-exclusive and the map is -1."""
    st_mp = np.asarray(st_mp).reshape(-1, 7)
    tst_mp = np.asarray(tst_mp).reshape(-1, 6)
    ntst = len(tst_mp)
    tst2st = -np.ones((ntst, 2), np.int64)
    excl = np.array(excl, dtype=np.int64).reshape(ntst, 2)
    for t in range(ntst):
        for q in range(2):
            b = int(tst_mp[t, q])
            o = b - 3 if side[t][q] == 1 else b + 3
            hit = np.nonzero(((st_mp[:, 0] == b) & (st_mp[:, 1] == o)) | ((st_mp[:, 0] == o) & (st_mp[:, 1] == b)))[0]
            if hit.size:
                tst2st[t, q] = int(hit[0])
            else:
                excl[t, q] = 0
    return tst2st, excl


def write_ninfo(system: System, path):
    """Write the DT15 records of a system in the current format (all-in-one
    numbering; unit columns from the system's units).  Angles are written
    in degrees with 17 significant digits; stack U0 is written as its
    (h, s, Tm), which read_ninfo turns back into U0 at the system
    temperature.  Tertiary stack types carry the side and exclusivity of
    each base (the sign gives the side, 2 marks exclusive)."""
    s = system
    u = lambda i: int(s.unit[i]) + 1  # noqa: E731
    ul = lambda i: i - _unit_start(s, int(s.unit[i])) + 1  # noqa: E731
    names = bead_names(s)
    deg_a = lambda r: r * 180.0 / T.F_PI  # noqa: E731
    out = ["<<<< native bond length\n"]
    for k, ((a, b), nat, coef) in enumerate(zip(s.bd_mp, s.bd_nat, s.bd_coef)):
        out.append("bond %6d %6d %6d %6d %6d %6d %6d %r 1.0 1.0 %r\n"
                   % (k + 1, u(a), u(b), a + 1, b + 1, ul(a), ul(b), float(nat), float(coef)))
    out.append(">>>>\n\n<<<< native bond angles\n")
    for k, ((a, b, c), nat, coef) in enumerate(zip(s.ba_mp, s.ba_nat, s.ba_coef)):
        out.append("angl %6d %6d %6d %6d %6d %6d %6d %6d %6d %r 1.0 1.0 %r\n"
                   % (k + 1, u(a), u(c), a + 1, b + 1, c + 1, ul(a), ul(b), ul(c), deg_a(float(nat)), float(coef)))
    out.append(">>>>\n\n<<<< native basestack\n")
    for k in range(len(s.st_mp)):
        m = s.st_mp[k]
        h, sv, tm = s.st_hsTm[k]
        out.append("bs-dist %6d %6d %6d %6d %6d %6d %6d %r %r %r %r %r %s-%s\n"
                   % (k + 1, u(m[0]), u(m[1]), m[0] + 1, m[1] + 1, ul(m[0]), ul(m[1]), float(h), float(sv),
                      float(tm), float(s.st_nat[k, 0]), float(s.st_coef[k, 1]), names[m[0]], names[m[1]]))
    for k in range(len(s.st_mp)):
        m = s.st_mp[k]
        for idih, (cols, typ) in enumerate((((2, 3, 4, 5), "PSPS"), ((3, 4, 5, 6), "SPSP"))):
            q = [int(m[c]) for c in cols]
            out.append("bs-dihd %6d %6d %6d %6d %s %s %r %r %s\n"
                       % (k + 1, idih + 1, u(q[0]), u(q[3]), " ".join("%6d" % (x + 1) for x in q),
                          " ".join("%6d" % ul(x) for x in q), deg_a(float(s.st_nat[k, 1 + idih])),
                          float(s.st_coef[k, 2 + idih]), typ))
    out.append(">>>>\n\n<<<< native hbond\n")
    site_atom = _site_atoms(names)
    for k in range(len(s.hb_mp)):
        m = [int(x) for x in s.hb_mp[k]]
        atoms = []
        for q in range(3):
            a, b = int(s.hb_site[k, 2 * q]), int(s.hb_site[k, 2 * q + 1])
            if a < 0:
                break
            atoms.append("%s %s" % (site_atom[a], site_atom[b]))
        out.append("hb-dist %6d %6d %6d %6d %6d %6d %6d %r %r %r %s %d %s\n"
                   % (k + 1, u(m[0]), u(m[1]), m[0] + 1, m[1] + 1, ul(m[0]), ul(m[1]),
                      float(s.hb_coef[k, 0]), float(s.hb_nat[k, 0]), float(s.hb_coef[k, 1]),
                      "T" if s.hb_tertiary[k] else "S", len(atoms), " ".join(atoms)))
        for iba, q in enumerate(((m[2], m[0], m[1]), (m[0], m[1], m[3]))):
            out.append("hb-angl %6d %6d %6d %6d %s %s %r %r\n"
                       % (k + 1, iba + 1, u(q[0]), u(q[2]), " ".join("%6d" % (x + 1) for x in q),
                          " ".join("%6d" % ul(x) for x in q), deg_a(float(s.hb_nat[k, 1 + iba])),
                          float(s.hb_coef[k, 2 + iba])))
        for idih, q in enumerate(((m[2], m[0], m[1], m[3]), (m[4], m[2], m[0], m[1]), (m[0], m[1], m[3], m[5]))):
            out.append("hb-dihd %6d %6d %6d %6d %s %s %r %r\n"
                       % (k + 1, idih + 1, u(q[0]), u(q[3]), " ".join("%6d" % (x + 1) for x in q),
                          " ".join("%6d" % ul(x) for x in q), deg_a(float(s.hb_nat[k, 3 + idih])),
                          float(s.hb_coef[k, 4 + idih])))
    out.append(">>>>\n\n<<<< native tertiary stack\n")
    for k in range(len(s.tst_mp)):
        m = [int(x) for x in s.tst_mp[k]]
        types = []
        for q in range(2):
            st = int(s.tst2st[k, q])
            if st >= 0:
                o = int(s.st_mp[st, 1]) if int(s.st_mp[st, 0]) == m[q] else int(s.st_mp[st, 0])
                sd = -1 if o < m[q] else 1
            else:
                sd = 1
            types.append(sd * (2 if s.tst_excl[k, q] else 1))
        out.append("tbs-dist %6d %6d %6d %6d %6d %6d %6d %r %r %r %+d %+d\n"
                   % (k + 1, u(m[0]), u(m[1]), m[0] + 1, m[1] + 1, ul(m[0]), ul(m[1]), float(s.tst_coef[k, 0]),
                      float(s.tst_nat[k, 0]), float(s.tst_coef[k, 1]), types[0], types[1]))
        for iba, q in enumerate(((m[2], m[0], m[1]), (m[0], m[1], m[3]))):
            out.append("tbs-angl %6d %6d %6d %6d %s %s %r %r\n"
                       % (k + 1, iba + 1, u(q[0]), u(q[2]), " ".join("%6d" % (x + 1) for x in q),
                          " ".join("%6d" % ul(x) for x in q), deg_a(float(s.tst_nat[k, 1 + iba])),
                          float(s.tst_coef[k, 2 + iba])))
        for idih, q in enumerate(((m[2], m[0], m[1], m[3]), (m[4], m[2], m[0], m[1]), (m[0], m[1], m[3], m[5]))):
            out.append("tbs-dihd %6d %6d %6d %6d %s %s %r %r\n"
                       % (k + 1, idih + 1, u(q[0]), u(q[3]), " ".join("%6d" % (x + 1) for x in q),
                          " ".join("%6d" % ul(x) for x in q), deg_a(float(s.tst_nat[k, 3 + idih])),
                          float(s.tst_coef[k, 4 + idih])))
    out.append(">>>>\n")
    with open(path, "w") as fh:
        fh.writelines(out)


def _unit_start(s: System, u: int) -> int:
    return int(np.nonzero(np.asarray(s.unit) == u)[0][0])


_SITE_ATOMS = {"S": ("O2'", "O4'", "O3'", "O5'"), "P": ("OP1", "OP2"), "A": ("N1", "N3", "N6", "N7"),
               "C": ("O2", "N3", "N4"), "G": ("N1", "N2", "N3", "O6", "N7"), "U": ("O2", "N3", "O4")}


def _site_atoms(names):
    """Atom name of every hbond site (the inverse of site_local)."""
    out = []
    for nm in names:
        if nm in _SITE_ATOMS:
            out.extend(_SITE_ATOMS[nm])
    return out
