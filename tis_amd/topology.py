"""DT15 topology and parameter derivation for RNA chains and ions.

The reference reads DT15 topology only from ``.ninfo`` files
(src/setp_nativestruct.F90:38-40, record formats src/read_nativeinfo.F90:121-705)
and has no sequence-to-topology generator.  This module writes the same
records directly from a sequence: bonds, angles, consecutive stacks and
secondary hydrogen bonds of a designed hairpin, with the A-form reference
geometry and DT15 coefficients of the reference's parameter files
(para/rna.para:236-416, para/general.para, para/electrostatic.para).

Temperature/salt dependent coefficients restate src/simu_ele_set.F90:5-218
(dielectric constant, Bjerrum length, Manning charge, DH prefactor, Debye
length) and src/simu_set_dtrna.F90:40-58 (stack U0(T), beta_hb).  Friction and
mass restate src/setp_mass_fric.F90:33-96.
"""
from __future__ import annotations

import math
from dataclasses import replace

import numpy as np

from .sysdesc import (CLASS_ION, CLASS_RNA, ELE_DH, ELE_EWALD, ELE_NONE, MP_BASE,
                      MP_ION_CA2, MP_ION_CL, MP_ION_K, MP_ION_MG, MP_ION_NA, MP_PHOS,
                      MP_SUGAR, Params, System)

# ---- physical constants, src/const_physical.F90:9-22 ----
F_PI = 3.14159265358979323846264338
EPSI_0 = 8.854187817e-12
ELE = 1.6021766208e-19
BOLTZ_J = 1.38064852e-23
N_AVO = 6.022140857e23
KCAL2JOUL = 4184.0
JOUL2KCAL = 1.0 / KCAL2JOUL
JOUL2KCAL_MOL = JOUL2KCAL * N_AVO
BOLTZ_KCAL_MOL = BOLTZ_J * JOUL2KCAL_MOL

# ---- para/rna.para <<<< A-form_RNA (Angstrom, degrees) ----
BOND_NAT = {"SP": 3.8157, "PS": 4.6010, "SA": 4.8515, "SU": 4.2733, "SG": 4.9659, "SC": 4.2738}
ANGL_NAT = {"PSP": 82.735, "SPS": 87.410, "PSA": 97.569, "PSU": 90.155, "PSG": 101.356,
            "PSC": 90.545, "ASP": 110.346, "USP": 112.661, "GSP": 109.721, "CSP": 112.615}
DIHD_NAT = {"PSPS": -148.215, "SPSP": 175.975}
STACK_DIST = {"AA": 4.1806530, "AC": 3.8260185, "AG": 4.4255305, "AU": 3.8260185,
              "CA": 4.7010580, "CC": 4.2500910, "CG": 4.9790760, "CU": 4.2273615,
              "GA": 4.0128560, "GC": 3.6784360, "GG": 4.2427250, "GU": 3.6616930,
              "UA": 4.7010580, "UC": 4.2679180, "UG": 4.9977560, "UU": 4.2453650}
HB_DIST = {"AU": 5.8815, "GC": 5.6550}
HB_ANGL = {"SA-U": 156.320, "SU-A": 143.910, "SG-C": 161.746, "SC-G": 142.306}
HB_DIHD = {"SA-US": 71.958, "SG-CS": 79.653, "PSA-U": 54.689, "PSU-A": 67.305,
           "PSG-C": 43.654, "PSC-G": 69.752}
# ---- para/rna.para <<<< DT15_stack_param: h, s, Tm ----
STACK_HSTM = {"AA": (3.99578, -0.319, 299.058), "AC": (3.96314, -0.319, 299.058),
              "AG": (4.75714, 5.301, 341.349), "AU": (3.96314, -0.319, 299.058),
              "CA": (3.92173, -0.319, 299.058), "CC": (3.65726, -1.567, 285.968),
              "CG": (4.23498, 0.774, 315.673), "CU": (3.62896, -1.567, 285.968),
              "GA": (4.71668, 5.301, 341.349), "GC": (4.71997, 4.370, 343.363),
              "GG": (5.19674, 7.346, 366.523), "GU": (4.61901, 2.924, 338.329),
              "UA": (3.92173, -0.319, 299.058), "UC": (3.62234, -1.567, 285.968),
              "UG": (4.67613, 2.924, 338.329), "UU": (2.99569, -3.563, 251.733)}
# ---- para/rna.para <<<< para_DT15_rna ----
BD_COEF = {"PS": 23.0, "SB": 10.0, "SP": 64.0}
BA_COEF = {"PSB": 5.0, "PSP": 20.0, "BSP": 5.0, "SPS": 20.0}
ST_DIST_K, ST_DIH_K = 1.4, 4.0
HB_DIST_K, HB_ANGL_K, HB_DIH_HBOND_K, HB_DIH_CHAIN_K = 5.0, 1.5, 0.15, 0.15
HB_U0_PER_HB = -2.7     # -8.1 for G-C (3 hbonds), -5.4 for A-U (2), ninfo/bwyv19.ninfo
# ---- para/rna.para <<<< DT15_exv_param: R, epsilon ----
EXV_PARAM = {"P": (2.1, 0.2), "S": (2.9, 0.2), "A": (2.8, 0.2), "G": (3.0, 0.2),
             "C": (2.7, 0.2), "U": (2.7, 0.2), "Mg": (0.7926, 0.894700),
             "Ca": (1.7131, 0.459789), "Cl": (1.9480, 0.265000), "K": (2.6580, 0.000328),
             "Na": (1.8680, 0.002770)}
# ---- para/general.para: mass, Stokes radius ----
MASS = {"P": 62.972561, "S": 131.10668, "A": 134.11876, "G": 150.11816, "C": 110.09406,
        "U": 111.07882, "Mg": 24.3050, "Cl": 35.4500, "K": 39.1000, "Na": 22.9900, "Ca": 40.078}
STOKES_R = {"P": 3.1, "S": 4.0, "A": 3.9, "G": 4.1, "C": 3.8, "U": 3.8,
            "Mg": 4.28, "Cl": 3.32, "K": 3.31, "Na": 3.58, "Ca": 4.12}
ION_CHARGE = {"Mg": 2.0, "Ca": 2.0, "K": 1.0, "Na": 1.0, "Cl": -1.0}
ION_MPTYPE = {"Mg": MP_ION_MG, "Ca": MP_ION_CA2, "Na": MP_ION_NA, "K": MP_ION_K, "Cl": MP_ION_CL}
LENGTH_PER_UNIT_P = 4.38178046   # para/electrostatic.para:42
NEIGH_MARGIN = 8.0               # para/general.para
# hydrogen-bond sites per bead type and their valences, src/setp_dtrna15.F90:41-161
HBSITE_VALENCE = {"S": (2, 1, 1, 1), "P": (1, 1), "A": (1, 1, 2, 1), "C": (2, 1, 2),
                  "G": (1, 2, 1, 2, 1), "U": (2, 1, 1)}
# local site numbers, src/read_nativeinfo.F90:1286-1336
_SITE_LOCAL = {("A", "N1"): 1, ("A", "N3"): 2, ("A", "N6"): 3, ("A", "N7"): 4,
               ("C", "O2"): 1, ("C", "N3"): 2, ("C", "N4"): 3,
               ("G", "N1"): 1, ("G", "N2"): 2, ("G", "N3"): 3, ("G", "O6"): 4, ("G", "N7"): 5,
               ("U", "O2"): 1, ("U", "N3"): 2, ("U", "O4"): 3}
# atom pairs of canonical base pairs (first atom on the first base), ninfo/bwyv19.ninfo
_PAIR_ATOMS = {("G", "C"): (("N1", "N3"), ("N2", "O2"), ("O6", "N4")),
               ("A", "U"): (("N1", "N3"), ("N6", "O4"))}
COMPLEMENT = {"A": "U", "U": "A", "G": "C", "C": "G"}
DE_MAX = 50.0


def deg2rad_angl(deg: float) -> float:
    """Bond angles: (ba / 180.0) * F_PI, src/read_nativeinfo.F90:231."""
    return (deg / 180.0) * F_PI


def deg2rad_hb(deg: float) -> float:
    """Dihedrals and hbond angles: x * F_PI / 180.0, src/read_nativeinfo.F90:761,843,907."""
    return deg * F_PI / 180.0


# ---------------------------------------------------------------------------
# temperature / salt dependent coefficients
def dielectric(tempk: float) -> float:
    """i_diele = 1 water dielectric, src/simu_ele_set.F90:31-63."""
    MM_A, MM_B, MM_C, MM_D = 87.740, -0.4008, 9.398e-4, -1.410e-6
    Tc = tempk - 273.15
    return MM_A + MM_B * Tc + MM_C * Tc * Tc + MM_D * Tc * Tc * Tc


def ele_params(tempk: float, ionic_strength: float):
    """Return (coef, cdist, Zp): DH prefactor 332.06/eps [kcal/mol A], Debye
    length [A] and the reduced phosphate charge b/l_B (i_charge = 1),
    src/simu_ele_set.F90:117-218 (model 2015, i_temp_independent = 0)."""
    ek = dielectric(tempk)
    lb = ELE * ELE / (4.0 * F_PI * EPSI_0 * ek * BOLTZ_J * tempk) * 1.0e10
    zp = LENGTH_PER_UNIT_P / lb
    coef = JOUL2KCAL_MOL * 1.0e10 * ELE ** 2 / (4.0 * F_PI * EPSI_0 * ek)
    cdist = 1.0e10 * math.sqrt((1.0e-3 * EPSI_0 * ek * BOLTZ_J) / (2.0 * N_AVO * ELE ** 2)) \
        * math.sqrt(tempk / ionic_strength)
    return coef, cdist, zp


def stack_u0(h: float, s: float, tm: float, tempk: float) -> float:
    """U0 = -h + kB (T - Tm) s, src/simu_set_dtrna.F90:40-52."""
    return -h + BOLTZ_KCAL_MOL * (tempk - tm) * s


def beta_hb(tempk: float) -> float:
    """beta_hbond15, src/simu_set_dtrna.F90:58."""
    return 1.0 / (tempk * BOLTZ_KCAL_MOL)


def langevin_visc(viscosity_pa_s: float) -> float:
    """Viscosity in (Da)(A^-1)(tau^-1), src/setp_mass_fric.F90:47-50."""
    return viscosity_pa_s * math.sqrt(1.0e3 * JOUL2KCAL) * N_AVO * 1.0e-20


# ---------------------------------------------------------------------------
def _bead_layout(seq: str):
    """Bead names of a 5'-S linear chain: S1 B1 P2 S2 B2 ... (N = 3n-1)."""
    names, idx = [], []
    for k, b in enumerate(seq):
        if k > 0:
            idx.append(("P", k)); names.append("P")
        idx.append(("S", k)); names.append("S")
        idx.append(("B", k)); names.append(b)
    return names, idx


class _Chain:
    def __init__(self, seq: str, base_index: int):
        self.seq = seq
        self.names, idx = _bead_layout(seq)
        self.pos = {key: base_index + i for i, key in enumerate(idx)}

    def P(self, k):
        return self.pos.get(("P", k), -1)

    def S(self, k):
        return self.pos[("S", k)]

    def B(self, k):
        return self.pos[("B", k)]


def hairpin_pairs(n: int, stem5: int, stem_len: int, loop: int):
    """Base-pair list (i, j) of a hairpin whose 5' stem starts at stem5."""
    j0 = stem5 + 2 * stem_len + loop - 1
    return [(stem5 + k, j0 - k) for k in range(stem_len)]


def hbsite_layout(names):
    """Hydrogen-bond sites of the RNA beads in bead order: {bead: first site}
    and the site valences, src/setp_dtrna15.F90:41-75,114-161."""
    offset, valence, site = {}, [], 0
    for i, nm in enumerate(names):
        if nm not in HBSITE_VALENCE:
            continue
        offset[i] = site
        valence.extend(HBSITE_VALENCE[nm])
        site += len(HBSITE_VALENCE[nm])
    return offset, valence


def assemble_system(names, unit, unit_class, topo, tempk=310.0, ionic_strength=0.150,
                    viscosity=1.0e-5, ele_form=ELE_DH, cutoff_ele=99.0, dh_cutoff_type=0,
                    neigh_margin=NEIGH_MARGIN, box=None, ewld_alpha=0.0, ewld_hmax=0,
                    phos_charge=None, seq="", name=""):
    """A System from bead names ('P', 'S', 'A'/'C'/'G'/'U' for bases, ion
    species 'Mg', 'K', ...), their units, the class of each unit and the
    topology records in ``topo`` (bd_mp/bd_nat/bd_coef, ba_*, st_* with
    st_hsTm, hb_* with hb_site/hb_tertiary, tst_* with tst2st/tst_excl).
    Per-bead excluded-volume, mass and Stokes parameters follow
    src/setp_dtrna15.F90:77-161 and src/setp_mass_fric.F90:33-96; the
    phosphate charge is -Zp(T, I) (src/simu_ele_set.F90:155-185, i_charge = 1)
    unless phos_charge is given; list radii per unit pair follow
    src/mloop_neigh_dist.F90:25-60,149."""
    coef, cdist, zp = ele_params(tempk, ionic_strength)
    visc = langevin_visc(viscosity)
    mptype, iclass, exv_rad, exv_eps, charge, mass, radius = [], [], [], [], [], [], []
    for nm in names:
        if nm in ION_MPTYPE:
            mptype.append(ION_MPTYPE[nm]); iclass.append(CLASS_ION)
            charge.append(ION_CHARGE[nm])
        else:
            mptype.append(MP_PHOS if nm == "P" else (MP_SUGAR if nm == "S" else MP_BASE))
            iclass.append(CLASS_RNA)
            charge.append((-zp if phos_charge is None else float(phos_charge)) if nm == "P" else 0.0)
        r, e = EXV_PARAM[nm]
        exv_rad.append(r); exv_eps.append(math.sqrt(e))
        mass.append(MASS[nm]); radius.append(STOKES_R[nm])
    _, hbsite_valence = hbsite_layout(names)
    exv_rad_a = np.asarray(exv_rad)
    unit_a = np.asarray(unit)
    nunit = len(unit_class)
    rl2 = np.zeros((nunit, nunit))
    p = Params()
    for a in range(nunit):
        for b in range(nunit):
            if unit_class[a] == CLASS_RNA and unit_class[b] == CLASS_RNA:
                mc = max(0.0, p.exv_dist, p.exv_dist_PS)
            else:
                ra = exv_rad_a[unit_a == a]
                rb = exv_rad_a[unit_a == b]
                mc = max(0.0, float((ra[:, None] + rb[None, :]).max()))
            rl2[a, b] = (mc + neigh_margin) ** 2
    mass_a = np.asarray(mass)
    fric = 6.0 * F_PI * visc * np.asarray(radius) / mass_a
    params = replace(Params(), neigh_margin=neigh_margin, ele_form=ele_form,
                     dh_cutoff_type=dh_cutoff_type, cutoff_ele=cutoff_ele,
                     ele_coef=coef, ele_cdist=cdist, beta_hb=beta_hb(tempk),
                     ewld_alpha=ewld_alpha, ewld_hmax=ewld_hmax,
                     periodic=1 if box is not None else 0,
                     box=tuple(box) if box is not None else (0.0, 0.0, 0.0))
    nh = len(topo.get("hb_mp", []))
    extra = {k: topo[k] for k in ("tst_mp", "tst_nat", "tst_coef", "tst2st", "tst_excl") if k in topo}
    return System(unit=unit_a, mptype=mptype, iclass=iclass, exv_rad=exv_rad_a,
                  exv_eps=exv_eps, charge=charge, rlist2_unit=rl2,
                  bd_mp=topo["bd_mp"], bd_nat=topo["bd_nat"], bd_coef=topo["bd_coef"],
                  ba_mp=topo["ba_mp"], ba_nat=topo["ba_nat"], ba_coef=topo["ba_coef"],
                  st_mp=topo["st_mp"], st_nat=topo["st_nat"], st_coef=topo["st_coef"],
                  st_hsTm=topo["st_hsTm"],
                  hb_mp=topo["hb_mp"], hb_nat=topo["hb_nat"], hb_coef=topo["hb_coef"],
                  hb_site=topo["hb_site"],
                  hb_tertiary=topo.get("hb_tertiary", np.zeros(nh, np.int32)),
                  hbsite_valence=hbsite_valence,
                  mass=mass_a, fric=fric, radius=radius, tempk=tempk, params=params,
                  seq=seq, name=name, **extra)


def build_rna_system(chains, tempk=310.0, ionic_strength=0.150, viscosity=1.0e-5,
                     ele_form=ELE_DH, cutoff_ele=99.0, dh_cutoff_type=0,
                     neigh_margin=NEIGH_MARGIN, ions=None, box=None,
                     ewld_alpha=0.0, ewld_hmax=0, all_pair_hbonds=False, phos_charge=None, name=""):
    """Build one DT15 system.

    chains: list of (sequence, base_pairs) with base_pairs [(i, j), ...]
      (0-based nucleotide indices, i < j, canonical pairs only).
    ions:   list of ion species names ('Mg', 'K', 'Cl', ...) forming one unit.
    all_pair_hbonds: make every canonical (i, j) with j - i >= 4 a candidate
      (a tertiary-rich candidate set, as in ninfo/azo.hb-dist.ninfo), instead
      of the designed stem pairs only.
    phos_charge: None for the reduced phosphate charge -Zp(T, I) of
      simu_ele_set.F90:155-185 (i_charge = 1), else the fixed value (e.g.
      -1.0 with explicit ions).
    """
    unit, names = [], []
    bd_mp, bd_nat, bd_coef = [], [], []
    ba_mp, ba_nat, ba_coef = [], [], []
    st_mp, st_nat, st_coef, st_hsTm = [], [], [], []
    hb_mp, hb_nat, hb_coef, hb_site = [], [], [], []
    cobjs = []
    base = 0
    for u, (seq, pairs) in enumerate(chains):
        ch = _Chain(seq, base)
        cobjs.append((ch, pairs))
        for nm in ch.names:
            unit.append(u)
            names.append(nm)
        base += len(ch.names)
    hbsite_offset, _ = hbsite_layout(names)
    for ch, pairs in cobjs:
        seq, n = ch.seq, len(ch.seq)
        # bonds: PS, SB, SP per nucleotide
        for k in range(n):
            if k > 0:
                bd_mp.append((ch.P(k), ch.S(k))); bd_nat.append(BOND_NAT["PS"]); bd_coef.append(BD_COEF["PS"])
            bd_mp.append((ch.S(k), ch.B(k))); bd_nat.append(BOND_NAT["S" + seq[k]]); bd_coef.append(BD_COEF["SB"])
            if k < n - 1:
                bd_mp.append((ch.S(k), ch.P(k + 1))); bd_nat.append(BOND_NAT["SP"]); bd_coef.append(BD_COEF["SP"])
        # angles
        for k in range(n):
            b = seq[k]
            if k > 0:
                ba_mp.append((ch.P(k), ch.S(k), ch.B(k)))
                ba_nat.append(deg2rad_angl(ANGL_NAT["PS" + b])); ba_coef.append(BA_COEF["PSB"])
                if k < n - 1:
                    ba_mp.append((ch.P(k), ch.S(k), ch.P(k + 1)))
                    ba_nat.append(deg2rad_angl(ANGL_NAT["PSP"])); ba_coef.append(BA_COEF["PSP"])
            if k < n - 1:
                ba_mp.append((ch.B(k), ch.S(k), ch.P(k + 1)))
                ba_nat.append(deg2rad_angl(ANGL_NAT[b + "SP"])); ba_coef.append(BA_COEF["BSP"])
                ba_mp.append((ch.S(k), ch.P(k + 1), ch.S(k + 1)))
                ba_nat.append(deg2rad_angl(ANGL_NAT["SPS"])); ba_coef.append(BA_COEF["SPS"])
        # consecutive stacks, src/read_nativeinfo.F90:482-488 (needs P_k and P_k+2)
        for k in range(1, n - 2):
            pair = seq[k] + seq[k + 1]
            st_mp.append((ch.B(k), ch.B(k + 1), ch.P(k), ch.S(k), ch.P(k + 1), ch.S(k + 1), ch.P(k + 2)))
            st_nat.append((STACK_DIST[pair], deg2rad_hb(DIHD_NAT["PSPS"]), deg2rad_hb(DIHD_NAT["SPSP"])))
            h, s, tm = STACK_HSTM[pair]
            st_hsTm.append((h, s, tm))
            st_coef.append((stack_u0(h, s, tm, tempk), ST_DIST_K, ST_DIH_K, ST_DIH_K))
        # hydrogen bonds
        cand = list(pairs)
        if all_pair_hbonds:
            cand = [(i, j) for i in range(n - 1) for j in range(i + 4, n - 1)
                    if COMPLEMENT[seq[i]] == seq[j]]
        for (i, j) in cand:
            if i > j:
                i, j = j, i
            if not (0 <= i and j < n - 1):
                raise ValueError("hbond partners need a 3' phosphate (j < n-1)")
            bi, bj = seq[i], seq[j]
            key = (bi, bj) if (bi, bj) in _PAIR_ATOMS else (bj, bi)
            atoms = _PAIR_ATOMS[key]
            if key != (bi, bj):
                atoms = tuple((b_, a_) for a_, b_ in atoms)
            hb_mp.append((ch.B(i), ch.B(j), ch.S(i), ch.S(j), ch.P(i + 1), ch.P(j + 1)))
            d0 = HB_DIST["AU" if "A" in (bi + bj) else "GC"]
            a1 = HB_ANGL["S%s-%s" % (bi, bj)]
            a2 = HB_ANGL["S%s-%s" % (bj, bi)]
            dh1 = HB_DIHD["SA-US" if "A" in (bi + bj) else "SG-CS"]
            dh2 = HB_DIHD["PS%s-%s" % (bi, bj)]
            dh3 = HB_DIHD["PS%s-%s" % (bj, bi)]
            hb_nat.append((d0, deg2rad_hb(a1), deg2rad_hb(a2), deg2rad_hb(dh1),
                           deg2rad_hb(dh2), deg2rad_hb(dh3)))
            hb_coef.append((HB_U0_PER_HB * len(atoms), HB_DIST_K, HB_ANGL_K, HB_ANGL_K,
                            HB_DIH_HBOND_K, HB_DIH_CHAIN_K, HB_DIH_CHAIN_K))
            sites = [-1] * 6
            for q, (ai, aj) in enumerate(atoms):
                sites[2 * q] = hbsite_offset[ch.B(i)] + _SITE_LOCAL[(bi, ai)] - 1
                sites[2 * q + 1] = hbsite_offset[ch.B(j)] + _SITE_LOCAL[(bj, aj)] - 1
            hb_site.append(sites)
    unit_class = [CLASS_RNA] * len(chains)
    if ions:
        u = len(chains)
        unit_class.append(CLASS_ION)
        for sp in ions:
            unit.append(u)
            names.append(sp)
    topo = dict(bd_mp=bd_mp, bd_nat=bd_nat, bd_coef=bd_coef, ba_mp=ba_mp, ba_nat=ba_nat,
                ba_coef=ba_coef, st_mp=st_mp, st_nat=st_nat, st_coef=st_coef, st_hsTm=st_hsTm,
                hb_mp=hb_mp, hb_nat=hb_nat, hb_coef=hb_coef, hb_site=hb_site)
    return assemble_system(names, unit, unit_class, topo, tempk=tempk,
                           ionic_strength=ionic_strength, viscosity=viscosity, ele_form=ele_form,
                           cutoff_ele=cutoff_ele, dh_cutoff_type=dh_cutoff_type,
                           neigh_margin=neigh_margin, box=box, ewld_alpha=ewld_alpha,
                           ewld_hmax=ewld_hmax, phos_charge=phos_charge,
                           seq="|".join(c[0] for c in chains), name=name)


# ---- para/rna.para:359-361 (st_nlocal_*), U0 of the tbs-dist records of ninfo/bwyv19.ninfo ----
TST_DIST_K, TST_ANGL_K, TST_DIH_K, TST_U0 = 5.0, 1.5, 0.15, -5.0


def _angle(a, b, c):
    u, v = a - b, c - b
    return math.acos(max(-1.0, min(1.0, float(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))))))


def _tstack_dihedrals(x, m):
    """The three dihedrals of a tertiary stack (B1, B2, S1, S2, P1, P2) in
    the convention of energy_dtrna_stack_nlocal.F90:163-212, so that a
    native taken here is a minimum of the energy."""
    v12, v13 = x[m[0]] - x[m[1]], x[m[0]] - x[m[2]]
    v53, v42, v46 = x[m[4]] - x[m[2]], x[m[3]] - x[m[1]], x[m[3]] - x[m[5]]
    c4212, c1213 = np.cross(v42, v12), np.cross(v12, v13)
    a12, a13, a42 = (float(np.linalg.norm(v)) for v in (v12, v13, v42))
    d1 = math.atan2(float(v42 @ c1213) * a12, float(c4212 @ c1213))
    d2 = math.atan2(float(v53 @ -c1213) * a13, float(np.cross(v53, v13) @ -c1213))
    n = np.cross(v42, v46)
    d3 = math.atan2(float(v12 @ n) * a42, float(-c4212 @ n))
    return d1, d2, d3


def add_tertiary_stacks(sys: System, xyz, stacks, u0=TST_U0, dist_shift=None, rng=None):
    """A copy of ``sys`` with tertiary (non-local) base stacks.

    stacks: [(B1, B2, type1, type2), ...] with B1, B2 0-based base beads and
    types as in tbs-dist records (read_nativeinfo.F90:624-705): the sign picks
    the side whose consecutive stack a base may lose (negative: the stack
    with the base three beads before, positive: after) and 2 marks the side
    exclusive, 1 not.  The beads of a stack are (B1, B2, S1 = B1-1,
    S2 = B2-1, P1 = B1+1, P2 = B2+1).  Native distance, angles and dihedrals
    are taken from ``xyz`` (minus dist_shift[k] when given, to place stack k
    that far from its native distance; with rng, N(0, 0.2) noise on the
    distance unless dist_shift is given, and N(0, 0.15) / N(0, 0.3) radians
    on the angles / dihedrals),
    coefficients from para/rna.para:359-361.  The tertiary-to-consecutive
    stack map follows src/mloop_dtrna.F90:73-136."""
    from .ninfo import tertiary_stack_map
    x = np.asarray(xyz, dtype=np.float64)
    tst_mp, tst_nat, tst_coef, side, excl = [], [], [], [], []
    for k, (b1, b2, t1, t2) in enumerate(stacks):
        if b1 > b2:
            b1, b2, t1, t2 = b2, b1, t2, t1
        m = (b1, b2, b1 - 1, b2 - 1, b1 + 1, b2 + 1)
        if sys.mptype[b1] != MP_BASE or sys.mptype[b2] != MP_BASE:
            raise ValueError("tertiary stacks join two base beads")
        d = float(np.linalg.norm(x[b1] - x[b2]))
        if dist_shift is not None:
            d -= float(dist_shift[k])
        elif rng is not None:
            d += float(rng.normal(0.0, 0.2))
        tst_mp.append(m)
        nat = [d, _angle(x[m[2]], x[m[0]], x[m[1]]), _angle(x[m[0]], x[m[1]], x[m[3]])]
        nat += list(_tstack_dihedrals(x, m))
        if rng is not None:
            nat[1:3] = [a + float(rng.normal(0.0, 0.15)) for a in nat[1:3]]
            nat[3:] = [a + float(rng.normal(0.0, 0.3)) for a in nat[3:]]
        tst_nat.append(tuple(nat))
        tst_coef.append((u0, TST_DIST_K, TST_ANGL_K, TST_ANGL_K, TST_DIH_K, TST_DIH_K, TST_DIH_K))
        side.append((1 if t1 < 0 else 2, 1 if t2 < 0 else 2))
        excl.append((1 if abs(t1) == 2 else 0, 1 if abs(t2) == 2 else 0))
    tst2st, excl_a = tertiary_stack_map(sys.st_mp, tst_mp, side, excl)
    return replace(sys, tst_mp=np.asarray(tst_mp), tst_nat=np.asarray(tst_nat),
                   tst_coef=np.asarray(tst_coef), tst2st=tst2st, tst_excl=excl_a)


def set_temperature(sys: System, tempk: float, ionic_strength: float) -> None:
    """Re-derive every T-dependent coefficient of a system in place
    (simu_para2.F90:8-108: simu_ele_set + simu_set_dtrna + charge refresh),
    keeping friction; the Langevin constants are re-derived by the engine."""
    coef, cdist, zp = ele_params(tempk, ionic_strength)
    sys.tempk = tempk
    sys.params.ele_coef, sys.params.ele_cdist = coef, cdist
    sys.params.beta_hb = beta_hb(tempk)
    phos = sys.mptype == MP_PHOS
    sys.charge[phos] = -zp
    for k in range(len(sys.st_coef)):
        h, s, tm = sys.st_hsTm[k]
        sys.st_coef[k, 0] = stack_u0(h, s, tm, tempk)


# ---------------------------------------------------------------------------
# coordinates
def helix_coords(sys: System, rng: np.random.Generator, jitter=0.5, rise=3.0,
                 twist_deg=36.0, r_s=8.0, r_b=3.5, chain_gap=60.0):
    """Extended single-strand helix per chain along z with N(0, jitter) noise."""
    xyz = np.zeros((sys.nmp, 3))
    tw = math.radians(twist_deg)
    units = sys.unit
    for u in np.unique(units):
        idx = np.nonzero((units == u) & (sys.iclass == CLASS_RNA))[0]
        if idx.size == 0:
            continue
        origin = np.array([u * chain_gap, 0.0, 0.0])
        k = -1
        for pos, i in enumerate(idx):
            t = sys.mptype[i]
            if t == MP_SUGAR:
                k += 1
                ang = k * tw
                xyz[i] = origin + [r_s * math.cos(ang), r_s * math.sin(ang), k * rise]
            elif t == MP_BASE:
                ang = k * tw
                xyz[i] = origin + [r_b * math.cos(ang), r_b * math.sin(ang), k * rise + 0.3]
            else:  # P between S_k and S_k+1, pushed outwards
                ang = (k + 0.45) * tw
                xyz[i] = origin + [(r_s + 2.0) * math.cos(ang), (r_s + 2.0) * math.sin(ang),
                                   (k + 0.45) * rise]
    return xyz + rng.normal(0.0, jitter, size=xyz.shape)


def compact_coords(sys: System, rng: np.random.Generator, box=40.0, step=4.2, min_dist=2.2,
                   tries=200):
    """Random self-avoiding walk per chain inside a cube: compact conformations in
    which excluded-volume, DH and hbond terms are all active (test inputs).  A
    step that leaves the cube is drawn again rather than clipped to the wall:
    clipping puts runs of beads on one edge line, where the stack dihedral of
    three collinear beads is undefined and the force NaN (in the reference's
    routines as well)."""
    xyz = np.zeros((sys.nmp, 3))
    placed = []
    for i in range(sys.nmp):
        for _ in range(tries):
            if i == 0 or sys.unit[i] != sys.unit[i - 1] or sys.iclass[i] != CLASS_RNA:
                cand = rng.uniform(-box / 2, box / 2, 3)
            else:
                ref = xyz[i - 1] if sys.mptype[i] != MP_PHOS else xyz[i - 2]
                if sys.mptype[i] == MP_BASE:
                    ref = xyz[i - 1]
                d = rng.normal(size=3)
                cand = ref + step * d / np.linalg.norm(d)
                if np.any(np.abs(cand) > box / 2):
                    continue
            if not placed or np.min(np.linalg.norm(np.asarray(placed) - cand, axis=1)) > min_dist:
                break
        xyz[i] = cand
        placed.append(cand)
    return xyz


def random_sequence(rng, n, gc_rich=False):
    if gc_rich:
        return "".join(rng.choice(list("GGGCCCAU"), size=n))
    return "".join(rng.choice(list("ACGU"), size=n))


def hairpin_design(rng, n, stem_len, loop):
    """Seeded hairpin: random flanks, a stem, a random loop, the complementary
    stem.  Returns (sequence, base_pairs)."""
    flank = n - 2 * stem_len - loop
    if flank < 2:
        raise ValueError("chain too short for the hairpin")
    f5 = int(rng.integers(1, flank))
    f3 = flank - f5
    stem = random_sequence(rng, stem_len, gc_rich=True)
    loop_s = random_sequence(rng, loop)
    comp = "".join(COMPLEMENT[c] for c in reversed(stem))
    seq = random_sequence(rng, f5) + stem + loop_s + comp + random_sequence(rng, f3)
    return seq, hairpin_pairs(len(seq), f5, stem_len, loop)


def config0_chain(seed=20261202):
    """64-nt hairpin: 16 random + 12-nt GC-rich stem + 8 loop + 12 complement + 16 random."""
    rng = np.random.default_rng(seed)
    stem = random_sequence(rng, 12, gc_rich=True)
    seq = (random_sequence(rng, 16) + stem + random_sequence(rng, 8)
           + "".join(COMPLEMENT[c] for c in reversed(stem)) + random_sequence(rng, 16))
    return seq, hairpin_pairs(64, 16, 12, 8)


def config2_ewald(seed=20261202, n=76, box=150.0, n_mg=10, n_k=305, n_cl=250, alpha=0.1,
                  rcut=30.0, hmax=10):
    """Explicit ions with Ewald: a random n-nt chain (phosphate charge -1) and
    n_mg Mg2+, n_k K+, n_cl Cl- placed uniformly at random without overlap in a
    periodic cube; full Coulomb by Ewald (alpha, real-space cutoff rcut,
    |h| <= hmax).  Returns (system, coordinates)."""
    rng = np.random.default_rng(seed)
    seq = random_sequence(rng, n)
    ions = ["Mg"] * n_mg + ["K"] * n_k + ["Cl"] * n_cl
    s = build_rna_system([(seq, [])], ions=ions, box=(box, box, box), ele_form=ELE_EWALD,
                         cutoff_ele=rcut, ewld_alpha=alpha, ewld_hmax=hmax, phos_charge=-1.0,
                         name=f"ewald{n}")
    x = compact_coords(s, rng, box=50.0, min_dist=3.0)
    rna = s.iclass == CLASS_RNA
    x[rna] -= x[rna].mean(axis=0)
    placed = [p for p in x[rna]]
    rad = s.exv_rad
    for i in np.nonzero(~rna)[0]:
        for _ in range(10000):
            c = rng.uniform(-box / 2, box / 2, 3)
            d = np.asarray(placed) - c
            d -= box * np.rint(d / box)
            if np.min(np.einsum("ij,ij->i", d, d)) >= (2.0 * rad[i] + 1.0) ** 2:
                break
        else:
            raise RuntimeError("could not place ion")
        x[i] = c
        placed.append(c)
    return s, x


def config1_lengths(seed=20261202, nchains=512):
    rng = np.random.default_rng(seed)
    n = np.rint(rng.lognormal(math.log(120.0), 0.7, size=nchains)).astype(int)
    return np.clip(n, 20, 600)
