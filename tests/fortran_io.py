"""Writes a DT15 system and coordinates in the stream layout that
tis_amd/fortran/tis_hip_check.F90 reads (little-endian int32 / float64)."""
import numpy as np


def write_system(path, s, xyz, nsteps=0, tstep=0.05):
    p = s.params
    hdr = np.array([s.nmp, s.nunit, len(s.bd_nat), len(s.ba_nat), len(s.st_mp), len(s.tst_mp), len(s.hb_mp),
                    len(s.hbsite_valence), p.ele_form, p.dh_cutoff_type, p.n_sep_P, p.n_sep_S, p.n_sep_B,
                    p.ewld_hmax, p.periodic, nsteps], dtype="<i4")
    par = np.array([p.exv_adjust, p.exv_dist, p.exv_dist_PS, p.exv_inf, p.hb_cutoff_dist, p.neigh_margin,
                    p.cutoff_ele, p.ele_coef, p.ele_cdist, p.ewld_alpha, p.beta_hb, *p.box, s.tempk, tstep, 0.0],
                   dtype="<f8")
    i4 = lambda a: np.ascontiguousarray(a, dtype="<i4").ravel()   # noqa: E731
    f8 = lambda a: np.ascontiguousarray(a, dtype="<f8").ravel()   # noqa: E731
    parts = [hdr, par, i4(s.unit), i4(s.mptype), i4(s.iclass), f8(s.exv_rad), f8(s.exv_eps), f8(s.charge),
             f8(s.rlist2_unit), i4(s.bd_mp), f8(s.bd_nat), f8(s.bd_coef), i4(s.ba_mp), f8(s.ba_nat),
             f8(s.ba_coef), i4(s.st_mp), f8(s.st_nat), f8(s.st_coef), i4(s.tst_mp), f8(s.tst_nat),
             f8(s.tst_coef), i4(s.tst2st), i4(s.tst_excl), i4(s.hb_mp), f8(s.hb_nat), f8(s.hb_coef),
             i4(s.hb_site), i4(s.hb_tertiary), i4(s.hbsite_valence), f8(s.mass), f8(s.fric), f8(s.radius),
             f8(xyz)]
    with open(path, "wb") as fh:
        for a in parts:
            fh.write(a.tobytes())


def read_result(path, nmp, nsteps=0):
    a = np.fromfile(path, dtype="<f8")
    f = a[:3 * nmp].reshape(nmp, 3)
    e = a[3 * nmp:3 * nmp + 25]
    x = a[3 * nmp + 25:].reshape(nmp, 3) if nsteps > 0 else None
    return f, e, x
