"""ctypes mirror of include/tis_system.h and a numpy-backed holder for one
DT15 system.

A ``System`` owns numpy arrays for every field of ``tis_system`` and builds
the C struct on demand (``System.cstruct()``), keeping the arrays alive for
as long as the struct is used.  The field meanings restate the reference's
module globals (src/var_struct.F90:17-288, src/var_setp.F90:190-200,563-576).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

# MPTYPE codes, src/const_index.F90:226-241
MP_PHOS, MP_BASE, MP_SUGAR = 2, 3, 4
MP_ION_MG, MP_ION_CA2, MP_ION_NA, MP_ION_K, MP_ION_CL = 5, 6, 7, 8, 9
# CLASS codes
CLASS_RNA, CLASS_ION = 2, 4
# E_TYPE slots (0-based), src/const_index.F90:99-137 (positional values 1..25)
E_TOTAL, E_VELO, E_BOND, E_BANGLE, E_ELE = 0, 1, 2, 3, 7
E_STACK, E_HBOND, E_EXV, E_TSTACK, E_THBOND = 15, 16, 17, 18, 19
E_MAX = 25
ELE_NONE, ELE_DH, ELE_COULOMB, ELE_EWALD = -1, 0, 1, 2

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class CParams(C.Structure):
    _fields_ = [
        ("exv_adjust", C.c_double), ("exv_dist", C.c_double),
        ("exv_dist_PS", C.c_double), ("exv_inf", C.c_double),
        ("n_sep_P", C.c_int32), ("n_sep_S", C.c_int32), ("n_sep_B", C.c_int32),
        ("hb_cutoff_dist", C.c_double), ("neigh_margin", C.c_double),
        ("ele_form", C.c_int32), ("dh_cutoff_type", C.c_int32),
        ("cutoff_ele", C.c_double), ("ele_coef", C.c_double),
        ("ele_cdist", C.c_double), ("ewld_alpha", C.c_double),
        ("ewld_hmax", C.c_int32), ("beta_hb", C.c_double),
        ("periodic", C.c_int32), ("box", C.c_double * 3),
    ]


class CSystem(C.Structure):
    _fields_ = [
        ("nmp", C.c_int32), ("nunit", C.c_int32),
        ("unit", _i32p), ("mptype", _i32p), ("iclass", _i32p),
        ("exv_rad", _f64p), ("exv_eps", _f64p), ("charge", _f64p),
        ("rlist2_unit", _f64p),
        ("nbd", C.c_int32), ("bd_mp", _i32p), ("bd_nat", _f64p), ("bd_coef", _f64p),
        ("nba", C.c_int32), ("ba_mp", _i32p), ("ba_nat", _f64p), ("ba_coef", _f64p),
        ("nst", C.c_int32), ("st_mp", _i32p), ("st_nat", _f64p), ("st_coef", _f64p),
        ("ntst", C.c_int32), ("tst_mp", _i32p), ("tst_nat", _f64p), ("tst_coef", _f64p),
        ("tst2st", _i32p), ("tst_excl", _i32p),
        ("nhb", C.c_int32), ("hb_mp", _i32p), ("hb_nat", _f64p), ("hb_coef", _f64p),
        ("hb_site", _i32p), ("hb_tertiary", _i32p),
        ("nhbsite", C.c_int32), ("hbsite_valence", _i32p),
        ("mass", _f64p), ("fric", _f64p), ("radius", _f64p),
        ("tempk", C.c_double),
        ("p", CParams),
    ]


@dataclass
class Params:
    exv_adjust: float = 1.5852
    exv_dist: float = 3.2
    exv_dist_PS: float = 3.2
    exv_inf: float = 0.2
    n_sep_P: int = 3
    n_sep_S: int = 3
    n_sep_B: int = 2
    hb_cutoff_dist: float = 2.0
    neigh_margin: float = 8.0
    ele_form: int = ELE_DH
    dh_cutoff_type: int = 0
    cutoff_ele: float = 99.0
    ele_coef: float = 0.0
    ele_cdist: float = 1.0
    ewld_alpha: float = 0.0
    ewld_hmax: int = 0
    beta_hb: float = 0.0
    periodic: int = 0
    box: tuple = (0.0, 0.0, 0.0)

    def cstruct(self) -> CParams:
        p = CParams()
        for name, _ in CParams._fields_:
            if name == "box":
                p.box = (C.c_double * 3)(*[float(b) for b in self.box])
            else:
                setattr(p, name, getattr(self, name))
        return p


def _i32(a, width=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    if width is not None and a.size == 0:
        a = np.zeros((0, width), dtype=np.int32)
    return a


def _f64(a, width=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if width is not None and a.size == 0:
        a = np.zeros((0, width), dtype=np.float64)
    return a


@dataclass
class System:
    """One DT15 system (one replica).  0-based indices; coordinates (N,3)."""
    unit: np.ndarray
    mptype: np.ndarray
    iclass: np.ndarray
    exv_rad: np.ndarray
    exv_eps: np.ndarray
    charge: np.ndarray
    rlist2_unit: np.ndarray          # (nunit, nunit)
    bd_mp: np.ndarray                # (nbd, 2)
    bd_nat: np.ndarray
    bd_coef: np.ndarray
    ba_mp: np.ndarray                # (nba, 3)
    ba_nat: np.ndarray
    ba_coef: np.ndarray
    st_mp: np.ndarray                # (nst, 7)
    st_nat: np.ndarray               # (nst, 3)
    st_coef: np.ndarray              # (nst, 4)
    st_hsTm: np.ndarray              # (nst, 3)  h, s, Tm for U0(T)
    hb_mp: np.ndarray                # (nhb, 6)
    hb_nat: np.ndarray               # (nhb, 6)
    hb_coef: np.ndarray              # (nhb, 7)
    hb_site: np.ndarray              # (nhb, 6)
    hb_tertiary: np.ndarray          # (nhb,)
    hbsite_valence: np.ndarray       # (nhbsite,)
    mass: np.ndarray
    fric: np.ndarray
    radius: np.ndarray
    tempk: float
    params: Params
    tst_mp: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.int32))
    tst_nat: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    tst_coef: np.ndarray = field(default_factory=lambda: np.zeros((0, 7)))
    tst2st: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    tst_excl: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    xyz: np.ndarray | None = None    # (N,3) starting coordinates
    seq: str = ""
    name: str = ""

    def __post_init__(self):
        for k in ("unit", "mptype", "iclass", "hb_tertiary", "hbsite_valence"):
            setattr(self, k, _i32(getattr(self, k)))
        for k, w in (("bd_mp", 2), ("ba_mp", 3), ("st_mp", 7), ("hb_mp", 6), ("hb_site", 6),
                     ("tst_mp", 6), ("tst2st", 2), ("tst_excl", 2)):
            setattr(self, k, _i32(getattr(self, k), w))
        for k in ("exv_rad", "exv_eps", "charge", "rlist2_unit", "bd_nat", "bd_coef",
                  "ba_nat", "ba_coef", "mass", "fric", "radius"):
            setattr(self, k, _f64(getattr(self, k)))
        for k, w in (("st_nat", 3), ("st_coef", 4), ("st_hsTm", 3), ("hb_nat", 6),
                     ("hb_coef", 7), ("tst_nat", 6), ("tst_coef", 7)):
            setattr(self, k, _f64(getattr(self, k), w))

    @property
    def nmp(self) -> int:
        return int(self.mptype.shape[0])

    @property
    def nunit(self) -> int:
        return int(self.rlist2_unit.shape[0])

    def cstruct(self) -> CSystem:
        def ip(a):
            return a.ctypes.data_as(_i32p)

        def fp(a):
            return a.ctypes.data_as(_f64p)

        s = CSystem()
        s.nmp, s.nunit = self.nmp, self.nunit
        s.unit, s.mptype, s.iclass = ip(self.unit), ip(self.mptype), ip(self.iclass)
        s.exv_rad, s.exv_eps, s.charge = fp(self.exv_rad), fp(self.exv_eps), fp(self.charge)
        s.rlist2_unit = fp(self.rlist2_unit)
        s.nbd, s.bd_mp, s.bd_nat, s.bd_coef = len(self.bd_nat), ip(self.bd_mp), fp(self.bd_nat), fp(self.bd_coef)
        s.nba, s.ba_mp, s.ba_nat, s.ba_coef = len(self.ba_nat), ip(self.ba_mp), fp(self.ba_nat), fp(self.ba_coef)
        s.nst, s.st_mp, s.st_nat, s.st_coef = len(self.st_mp), ip(self.st_mp), fp(self.st_nat), fp(self.st_coef)
        s.ntst, s.tst_mp, s.tst_nat, s.tst_coef = len(self.tst_mp), ip(self.tst_mp), fp(self.tst_nat), fp(self.tst_coef)
        s.tst2st, s.tst_excl = ip(self.tst2st), ip(self.tst_excl)
        s.nhb, s.hb_mp, s.hb_nat, s.hb_coef = len(self.hb_mp), ip(self.hb_mp), fp(self.hb_nat), fp(self.hb_coef)
        s.hb_site, s.hb_tertiary = ip(self.hb_site), ip(self.hb_tertiary)
        s.nhbsite, s.hbsite_valence = len(self.hbsite_valence), ip(self.hbsite_valence)
        s.mass, s.fric, s.radius = fp(self.mass), fp(self.fric), fp(self.radius)
        s.tempk = float(self.tempk)
        s.p = self.params.cstruct()
        return s
