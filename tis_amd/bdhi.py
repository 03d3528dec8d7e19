"""Brownian dynamics with hydrodynamic interaction: system description,
bead-spring chain builder and the ctypes mirror of include/tis_bdhi.h.

The chain builder produces what the reference reads from a .ninfo file for
this simulation type: FENE bonds (ifene2mp, fene_nat, coef_fene,
dist2_fene; force_fene.F90) and a static WCA pair list (iwca2mp, wca_nat2,
coef_wca; force_wca.F90), plus one hydrodynamic radius per bead.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from .topology import langevin_visc

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)

# i_hydro_tensor, const_index.F90:39-45
HI_RPY, HI_RPY_OVER, HI_ERMAK_OVER, HI_ZUK_RPY = 1, 2, 3, 4


class CBDSystem(C.Structure):
    _fields_ = [
        ("nmp", C.c_int32), ("radius", _f64p),
        ("nfene", C.c_int32), ("fene_mp", _i32p), ("fene_nat", _f64p), ("fene_coef", _f64p),
        ("fene_dist2", _f64p),
        ("nwca", C.c_int32), ("wca_mp", _i32p), ("wca_nat2", _f64p), ("wca_coef", _f64p),
        ("cutoff_wca", C.c_double), ("tempk", C.c_double), ("viscosity", C.c_double),
        ("hydro_tensor", C.c_int32),
    ]


@dataclass
class BDSystem:
    radius: np.ndarray               # (n,)
    fene_mp: np.ndarray              # (nfene, 2) 0-based
    fene_nat: np.ndarray
    fene_coef: np.ndarray
    fene_dist2: np.ndarray
    wca_mp: np.ndarray               # (nwca, 2)
    wca_nat2: np.ndarray
    wca_coef: np.ndarray             # (nwca, 2): repulsive, attractive
    cutoff_wca: float
    tempk: float
    viscosity: float                 # internal units (setp_mass_fric.F90:67)
    hydro_tensor: int = HI_ZUK_RPY

    def __post_init__(self):
        self.radius = np.ascontiguousarray(self.radius, dtype=np.float64)
        self.fene_mp = np.ascontiguousarray(np.asarray(self.fene_mp, dtype=np.int32).reshape(-1, 2))
        self.wca_mp = np.ascontiguousarray(np.asarray(self.wca_mp, dtype=np.int32).reshape(-1, 2))
        for k in ("fene_nat", "fene_coef", "fene_dist2", "wca_nat2"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float64))
        self.wca_coef = np.ascontiguousarray(np.asarray(self.wca_coef, dtype=np.float64).reshape(-1, 2))

    @property
    def nmp(self) -> int:
        return int(self.radius.shape[0])

    def cstruct(self) -> CBDSystem:
        def fp(a):
            return a.ctypes.data_as(_f64p)

        def ip(a):
            return a.ctypes.data_as(_i32p)
        s = CBDSystem()
        s.nmp = self.nmp
        s.radius = fp(self.radius)
        s.nfene, s.fene_mp = len(self.fene_nat), ip(self.fene_mp)
        s.fene_nat, s.fene_coef, s.fene_dist2 = fp(self.fene_nat), fp(self.fene_coef), fp(self.fene_dist2)
        s.nwca, s.wca_mp, s.wca_nat2, s.wca_coef = len(self.wca_nat2), ip(self.wca_mp), fp(self.wca_nat2), \
            fp(self.wca_coef)
        s.cutoff_wca, s.tempk, s.viscosity = self.cutoff_wca, self.tempk, self.viscosity
        s.hydro_tensor = self.hydro_tensor
        s._keep = self  # arrays stay alive with the struct
        return s


def bead_spring_chain(n, seed=0, radius=3.5, bond=7.0, fene_k=10.0, fene_r0=3.5, wca_eps=0.6,
                      wca_window=3.0, tempk=300.0, viscosity_pa_s=1.0e-5, hydro_tensor=HI_ZUK_RPY):
    """A bead-spring chain of n beads of hydrodynamic radius `radius` [A]:
    FENE bonds between neighbours (r0 = bond, maximum extension fene_r0),
    a static WCA list (r_min = bond) over pairs |i-j| >= 2 closer than
    wca_window * bond in the starting coordinates, purely repulsive
    (cutoff_wca = 1).  Start: a self-avoiding random walk of step `bond`.
    Returns (BDSystem, coordinates (n, 3))."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3))
    min2 = (0.9 * bond) ** 2
    for i in range(1, n):
        for _ in range(1000):
            u = rng.normal(size=3)
            c = x[i - 1] + bond * u / np.linalg.norm(u)
            if i < 2 or np.min(np.sum((x[:i - 1] - c) ** 2, axis=1)) >= min2:
                break
        x[i] = c
    fene = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    ii, jj = np.triu_indices(n, k=2)
    d2 = np.sum((x[ii] - x[jj]) ** 2, axis=1)
    keep = d2 < (wca_window * bond) ** 2
    wca = np.stack([ii[keep], jj[keep]], axis=1)
    nw = len(wca)
    sysd = BDSystem(radius=np.full(n, radius), fene_mp=fene, fene_nat=np.full(n - 1, bond),
                    fene_coef=np.full(n - 1, fene_k), fene_dist2=np.full(n - 1, fene_r0 ** 2),
                    wca_mp=wca, wca_nat2=np.full(nw, bond ** 2),
                    wca_coef=np.stack([np.full(nw, wca_eps), np.zeros(nw)], axis=1), cutoff_wca=1.0,
                    tempk=tempk, viscosity=langevin_visc(viscosity_pa_s), hydro_tensor=hydro_tensor)
    return sysd, x


def stokes_diffusion(s: BDSystem, i: int = 0) -> float:
    """kT / (6 pi eta a_i): the diagonal of D (A^2 per tau)."""
    from .topology import BOLTZ_KCAL_MOL
    return BOLTZ_KCAL_MOL * s.tempk / (6.0 * math.pi * s.viscosity * s.radius[i])
