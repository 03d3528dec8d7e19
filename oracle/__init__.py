"""Python handles on the CPU oracle (test infrastructure only).

``Oracle``  wraps liboracle.so, the FP64 C restatement (tis_oracle.c).
``RefLib``  wraps _ref/libtisref.so, the reference's own Fortran routines
            compiled from its sources (built only where they are present).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; it is what the HIP library is compared against, never
the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libtisref.so")
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/libtisref.so when the reference is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _lib():
    if not os.path.exists(LIB):
        build(ref=False)
    lib = C.CDLL(LIB)
    vp = C.c_void_p
    lib.oracle_ws_new.restype = vp
    lib.oracle_ws_new.argtypes = [vp]
    lib.oracle_ws_free.argtypes = [vp]
    lib.oracle_neighbor.argtypes = [vp, _f64p]
    lib.oracle_ws_lists.restype = C.c_int64
    lib.oracle_ws_lists.argtypes = [vp, C.c_int32, C.POINTER(_i32p), C.POINTER(_f64p)]
    lib.oracle_force_sumup.restype = C.c_int32
    lib.oracle_force_sumup.argtypes = [vp, _f64p, _f64p, C.c_uint64, C.c_uint32, C.c_uint64]
    lib.oracle_energy_sumup.argtypes = [vp, _f64p, _f64p, _f64p]
    lib.oracle_energy_unit.argtypes = [vp, _f64p, _f64p, C.c_int32, _f64p]
    lib.oracle_langevin_run.restype = C.c_int32
    lib.oracle_langevin_run.argtypes = [vp, C.c_double, C.c_int64, C.c_int32, _f64p, _f64p, _f64p,
                                        C.c_uint64, C.c_uint32, C.c_uint64, C.c_int32]
    lib.oracle_langevin_consts.argtypes = [vp, C.c_double, _f64p]
    lib.oracle_gauss3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _f64p]
    lib.oracle_uniform.restype = C.c_double
    lib.oracle_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]
    lib.oracle_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    for nm in ("oracle_force_bond", "oracle_force_bangle"):
        getattr(lib, nm).argtypes = [vp, _f64p, _f64p]
    lib.oracle_force_exv.restype = C.c_int32
    lib.oracle_force_exv.argtypes = [vp, _f64p, _i32p, C.c_int64, _f64p]
    lib.oracle_force_dh.restype = C.c_int32
    lib.oracle_force_dh.argtypes = [vp, _f64p, _i32p, _f64p, C.c_int64, _f64p]
    lib.oracle_force_stack.argtypes = [vp, _f64p, C.c_char_p, _f64p]
    lib.oracle_force_tstack.argtypes = [vp, _f64p, _f64p, C.c_char_p, _f64p]
    lib.oracle_hbond_eval.argtypes = [vp, _f64p, _f64p, _i32p, C.c_int32, _f64p, _f64p, C.c_char_p]
    lib.oracle_hbond_valence.restype = C.c_int32
    lib.oracle_hbond_valence.argtypes = [vp, _i32p, C.c_int32, _f64p, C.c_char_p,
                                         C.c_uint64, C.c_uint32, C.c_uint64]
    _u32p = C.POINTER(C.c_uint32)
    lib.oracle_hbond_valence_mt.restype = C.c_int32
    lib.oracle_hbond_valence_mt.argtypes = [vp, _i32p, C.c_int32, _f64p, C.c_char_p, _u32p,
                                            C.c_int64, C.POINTER(C.c_int64)]
    lib.oracle_force_sumup_mt.restype = C.c_int32
    lib.oracle_force_sumup_mt.argtypes = [vp, _f64p, _f64p, _u32p, C.c_int64, C.POINTER(C.c_int64)]
    lib.oracle_hbond_scatter.argtypes = [vp, _i32p, C.c_int32, _f64p, C.c_char_p, _f64p]
    lib.oracle_force_coulomb.argtypes = [vp, _f64p, _i32p, _f64p, C.c_int64, _f64p]
    lib.oracle_force_ewald.argtypes = [vp, _f64p, _i32p, _f64p, C.c_int64, _f64p]
    lib.oracle_energy_coulomb.restype = C.c_double
    lib.oracle_energy_coulomb.argtypes = [vp, _f64p, _i32p, _f64p, C.c_int64]
    lib.oracle_energy_ewald.restype = C.c_double
    lib.oracle_energy_ewald.argtypes = [vp, _f64p, _i32p, _f64p, C.c_int64, _f64p]
    lib.oracle_ewald_kvec.restype = C.c_int32
    lib.oracle_ewald_kvec.argtypes = [vp, _i32p, _f64p]
    lib.oracle_rpy_fill.argtypes = [C.c_int32, _f64p, _f64p, C.c_double, C.c_double, C.c_int32, _f64p]
    lib.oracle_cholesky.restype = C.c_int32
    lib.oracle_cholesky.argtypes = [C.c_int32, _f64p, _f64p]
    lib.oracle_force_fene.argtypes = [C.c_int32, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p]
    lib.oracle_force_wca.argtypes = [C.c_int32, _i32p, _f64p, _f64p, C.c_double, _f64p, _f64p]
    lib.oracle_bdhi_step.argtypes = [C.c_int32, _f64p, _f64p, _f64p, C.c_double, C.c_double, C.c_uint64,
                                     C.c_uint32, C.c_uint64, _f64p]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def fp(a):
    return a.ctypes.data_as(_f64p)


def ip(a):
    return a.ctypes.data_as(_i32p)


class Oracle:
    """FP64 restatement for one System (tis_amd.sysdesc.System)."""

    def __init__(self, system):
        self.sys = system
        self.cs = system.cstruct()
        self.L = lib()
        self.ws = self.L.oracle_ws_new(C.byref(self.cs))

    def __del__(self):
        try:
            self.L.oracle_ws_free(self.ws)
        except Exception:
            pass

    def neighbor(self, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        self.L.oracle_neighbor(self.ws, fp(x))

    def lists(self, which):
        p = _i32p()
        c = _f64p()
        n = self.L.oracle_ws_lists(self.ws, which, C.byref(p), C.byref(c))
        if which == 2:
            return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy()
        pairs = np.ctypeslib.as_array(p, shape=(max(n, 1) * 3,))[:3 * n].reshape(n, 3).copy()
        if which == 1:
            coef = np.ctypeslib.as_array(c, shape=(max(n, 1),))[:n].copy()
            return pairs, coef
        return pairs

    def force(self, xyz, seed=0, sys_id=0, step=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        f = np.zeros_like(x)
        ncl = self.L.oracle_force_sumup(self.ws, fp(x), fp(f), seed, sys_id, step)
        return f, ncl

    def force_mt(self, xyz, raw):
        """force_sumup whose valence pass draws the Mersenne-Twister words
        raw (uint32, in stream order).  Returns (f, words used)."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        r = np.ascontiguousarray(raw, dtype=np.uint32)
        f = np.zeros_like(x)
        used = C.c_int64()
        rc = self.L.oracle_force_sumup_mt(self.ws, fp(x), fp(f), r.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          len(r), C.byref(used))
        if rc < 0:
            raise ValueError("raw stream too short")
        return f, used.value

    def hbond_mt(self, xyz, raw):
        """The hbond term alone with the valence pass drawing raw (see
        force_mt).  Returns (f, status, ene, deletions, words used)."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        r = np.ascontiguousarray(raw, dtype=np.uint32)
        cs = C.byref(self.cs)
        hb = np.ascontiguousarray(self.lists(2), dtype=np.int32)
        nh = max(len(self.sys.hb_mp), 1)
        forh, ene, stat = np.zeros(18 * nh), np.zeros(nh), np.zeros(nh, np.uint8)
        f = np.zeros_like(x)
        self.L.oracle_hbond_eval(cs, fp(x), fp(x), ip(hb), len(hb), fp(forh), fp(ene),
                                 stat.ctypes.data_as(C.c_char_p))
        used = C.c_int64()
        nd = self.L.oracle_hbond_valence_mt(cs, ip(hb), len(hb), fp(ene), stat.ctypes.data_as(C.c_char_p),
                                            r.ctypes.data_as(C.POINTER(C.c_uint32)), len(r), C.byref(used))
        if nd < 0:
            raise ValueError("raw stream too short")
        self.L.oracle_hbond_scatter(cs, ip(hb), len(hb), fp(forh), stat.ctypes.data_as(C.c_char_p), fp(f))
        nreal = len(self.sys.hb_mp)
        return f, stat[:nreal], ene[:nreal], nd, used.value

    def energy(self, xyz, velo=None):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        e = np.zeros(25)
        v = None if velo is None else np.ascontiguousarray(velo, dtype=np.float64)
        self.L.oracle_energy_sumup(self.ws, fp(x), fp(v) if v is not None else None, fp(e))
        return e

    def energy_unit(self, xyz, velo=None, style=0):
        """energy_unit (nunit, nunit, 25) after the post-processing of
        i_output_energy_style = style (lists and hbond state of the last
        neighbor()/force() calls, like energy())."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        nu = self.sys.nunit
        eu = np.zeros(25 * nu * nu)
        v = None if velo is None else np.ascontiguousarray(velo, dtype=np.float64)
        self.L.oracle_energy_unit(self.ws, fp(x), fp(v) if v is not None else None, style, fp(eu))
        return eu.reshape(25, nu, nu).transpose(2, 1, 0)

    def langevin(self, xyz, velo, accel, nsteps, tstep=0.05, rebuild_every=10, seed=1,
                 sys_id=0, step0=0, init_accel=True):
        x = np.ascontiguousarray(xyz, dtype=np.float64).copy()
        v = np.ascontiguousarray(velo, dtype=np.float64).copy()
        a = np.ascontiguousarray(accel, dtype=np.float64).copy()
        nr = self.L.oracle_langevin_run(self.ws, tstep, nsteps, rebuild_every, fp(x), fp(v), fp(a),
                                        seed, sys_id, step0, 1 if init_accel else 0)
        return x, v, a, nr

    def langevin_consts(self, tstep=0.05):
        r = np.zeros((self.sys.nmp, 4))
        self.L.oracle_langevin_consts(C.byref(self.cs), tstep, fp(r))
        return r

    def term(self, name, xyz, pairs=None, coef=None):
        """One force term on its own (lists from the last neighbor())."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        f = np.zeros_like(x)
        cs = C.byref(self.cs)
        if name == "bond":
            self.L.oracle_force_bond(cs, fp(x), fp(f))
        elif name == "angle":
            self.L.oracle_force_bangle(cs, fp(x), fp(f))
        elif name == "exv":
            pr = np.ascontiguousarray(self.lists(0) if pairs is None else pairs, dtype=np.int32)
            self.L.oracle_force_exv(cs, fp(x), ip(pr), len(pr), fp(f))
        elif name == "dh":
            pr, cf = self.lists(1) if pairs is None else (pairs, coef)
            pr = np.ascontiguousarray(pr, dtype=np.int32)
            cf = np.ascontiguousarray(cf, dtype=np.float64)
            self.L.oracle_force_dh(cs, fp(x), ip(pr), fp(cf), len(pr), fp(f))
        elif name in ("coulomb", "ewald"):
            pr, cf = self.lists(1) if pairs is None else (pairs, coef)
            pr = np.ascontiguousarray(pr, dtype=np.int32)
            cf = np.ascontiguousarray(cf, dtype=np.float64)
            fn = self.L.oracle_force_coulomb if name == "coulomb" else self.L.oracle_force_ewald
            fn(cs, fp(x), ip(pr), fp(cf), len(pr), fp(f))
        elif name == "stack":
            st = np.ones(max(len(self.sys.st_mp), 1), np.uint8)
            self.L.oracle_force_tstack(cs, fp(x), fp(x), st.ctypes.data_as(C.c_char_p), fp(f))
            self.L.oracle_force_stack(cs, fp(x), st.ctypes.data_as(C.c_char_p), fp(f))
        elif name == "hbond":
            hb = np.ascontiguousarray(self.lists(2), dtype=np.int32)
            nh = max(len(self.sys.hb_mp), 1)
            forh = np.zeros(18 * nh)
            ene = np.zeros(nh)
            stat = np.zeros(nh, np.uint8)
            self.L.oracle_hbond_eval(cs, fp(x), fp(x), ip(hb), len(hb), fp(forh), fp(ene),
                                     stat.ctypes.data_as(C.c_char_p))
            self.L.oracle_hbond_valence(cs, ip(hb), len(hb), fp(ene),
                                        stat.ctypes.data_as(C.c_char_p), 0, 0, 0)
            self.L.oracle_hbond_scatter(cs, ip(hb), len(hb), fp(forh),
                                        stat.ctypes.data_as(C.c_char_p), fp(f))
            return f, stat, ene
        else:
            raise KeyError(name)
        return f


    def ewald_kvec(self):
        """Reciprocal vectors h (nk, 3) and coefficients f(h) of the system."""
        cs = C.byref(self.cs)
        nk = self.L.oracle_ewald_kvec(cs, None, None)
        h = np.zeros((max(nk, 1), 3), np.int32)
        c = np.zeros(max(nk, 1))
        self.L.oracle_ewald_kvec(cs, ip(h), fp(c))
        return h[:nk], c[:nk]

    def ewald_energy_parts(self, pxyz):
        """(Fourier, real, self) Ewald energies at wrapped coordinates pxyz
        with the ele list of the last neighbor()."""
        x = np.ascontiguousarray(pxyz, dtype=np.float64)
        pr, cf = self.lists(1)
        pr = np.ascontiguousarray(pr, dtype=np.int32)
        cf = np.ascontiguousarray(cf, dtype=np.float64)
        parts = np.zeros(3)
        self.L.oracle_energy_ewald(C.byref(self.cs), fp(x), ip(pr), fp(cf), len(pr), fp(parts))
        return parts


def gauss3(seed, sys_id, step, bead):
    g = np.zeros(3)
    lib().oracle_gauss3(seed, sys_id, step, bead, fp(g))
    return g


def uniform(seed, sys_id, step, stream, idx):
    return lib().oracle_uniform(seed, sys_id, step, stream, idx)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32(c, k, o)
    return list(o)


# ---------------------------------------------------------------------------
# Brownian dynamics with hydrodynamic interaction (BDSystem of tis_amd.bdhi)
def rpy_tensor(bd, xyz):
    """D (3n, 3n), simu_hydro_tensors.F90:17-302."""
    n = bd.nmp
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    D = np.zeros((3 * n, 3 * n), order="F")
    lib().oracle_rpy_fill(n, fp(x), fp(bd.radius), bd.tempk, bd.viscosity, bd.hydro_tensor,
                          D.ctypes.data_as(_f64p))
    return D


def cholesky(D):
    """Lower L with D = L L^T (simu_hydro_tensors.F90:304-324); raises on a
    non-positive pivot."""
    Df = np.asfortranarray(D, dtype=np.float64)
    L = np.zeros_like(Df, order="F")
    rc = lib().oracle_cholesky(Df.shape[0], Df.ctypes.data_as(_f64p), L.ctypes.data_as(_f64p))
    if rc:
        raise np.linalg.LinAlgError(f"non-positive pivot at row {rc - 1}")
    return L


def bd_forces(bd, xyz):
    """force_fene + force_wca (non-periodic)."""
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    f = np.zeros_like(x)
    L = lib()
    L.oracle_force_fene(len(bd.fene_nat), ip(bd.fene_mp), fp(bd.fene_nat), fp(bd.fene_coef),
                        fp(bd.fene_dist2), fp(x), fp(f))
    L.oracle_force_wca(len(bd.wca_nat2), ip(bd.wca_mp), fp(bd.wca_nat2), fp(bd.wca_coef), bd.cutoff_wca,
                       fp(x), fp(f))
    return f


def bdhi_run(bd, xyz, nsteps, tstep=0.05, refactor_every=1, seed=1, key=0, step0=0):
    """Ermak-McCammon steps (time_integral.F90:429-457); the tensors are
    rebuilt before the first step and every refactor_every steps."""
    x = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    D = L = None
    for t in range(nsteps):
        if t % refactor_every == 0:
            D = rpy_tensor(bd, x)
            L = cholesky(D)
        f = bd_forces(bd, x)
        lib().oracle_bdhi_step(bd.nmp, D.ctypes.data_as(_f64p), L.ctypes.data_as(_f64p), fp(f), bd.tempk, tstep,
                               seed, key, step0 + t + 1, fp(x))
    return x


# ---------------------------------------------------------------------------
def _on_big_stack(fn, *args):
    """Run fn(*args) on a thread with a 1 GiB stack: the reference's list
    builders keep MXMPELE * ncharge pair buffers as automatic (stack) arrays,
    which overflow the default 8 MiB stack for a few hundred charges."""
    import threading
    out, err = [], []

    def run():
        try:
            out.append(fn(*args))
        except BaseException as e:  # noqa: BLE001 - re-raised below
            err.append(e)
    old = threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if err:
        raise err[0]
    return out[0] if out else None


class RefLib:
    """The reference's own Fortran routines (one replica, one thread)."""

    TERMS = {"bond": 0, "angle": 1, "stack": 2, "stack_only": 3, "hbond": 4, "exv": 5, "dh": 6,
             "coulomb": 7, "ewald": 8}

    @staticmethod
    def available() -> bool:
        return os.path.exists(REFLIB)

    def __init__(self):
        self.L = C.CDLL(REFLIB)
        L = self.L
        L.tisref_load.argtypes = [C.c_void_p]
        L.tisref_neighbor.argtypes = [_f64p]
        L.tisref_move.argtypes = [_f64p]
        L.tisref_list_count.restype = C.c_int64
        L.tisref_list_count.argtypes = [C.c_int32]
        L.tisref_list_get.argtypes = [C.c_int32, _i32p, _f64p]
        L.tisref_force_term.argtypes = [C.c_int32, _f64p]
        L.tisref_hb_state.argtypes = [_i32p, _f64p]
        L.tisref_energy.argtypes = [_f64p, C.c_int32, _f64p]
        L.tisref_energy_unit.argtypes = [_f64p, C.c_int32, C.c_int32, _f64p]
        L.tisref_set_mass.argtypes = [_f64p]
        L.tisref_temperature.argtypes = [C.c_double, C.c_double, _f64p, _f64p, _f64p, _f64p,
                                         _f64p, _f64p]
        L.tisref_hydro.argtypes = [C.c_int32, _f64p, _f64p, C.c_double, C.c_double, C.c_int32, _f64p, _f64p]
        L.tisref_bd_forces.argtypes = [C.c_int32, _f64p, C.c_int32, _i32p, _f64p, _f64p, _f64p, C.c_int32,
                                       _i32p, _f64p, _f64p, C.c_double, _f64p]
        L.tisref_flush.argtypes = []
        L.tisref_mt_seed.argtypes = [C.c_int32]
        L.tisref_mt_raw.argtypes = [C.c_int32, C.c_int32, _i32p]
        self.sys = None

    def load(self, system):
        self.sys = system
        self.cs = system.cstruct()
        self.L.tisref_load(C.byref(self.cs))
        self._mass = np.ascontiguousarray(system.mass, dtype=np.float64)
        self.L.tisref_set_mass(fp(self._mass))

    def neighbor(self, xyz):
        self._x = np.ascontiguousarray(xyz, dtype=np.float64)
        _on_big_stack(self.L.tisref_neighbor, fp(self._x))

    def move(self, xyz):
        self._x = np.ascontiguousarray(xyz, dtype=np.float64)
        self.L.tisref_move(fp(self._x))

    def lists(self, which):
        n = self.L.tisref_list_count(which)
        if which == 2:
            p = np.zeros(max(n, 1), np.int32)
            self.L.tisref_list_get(which, ip(p), fp(np.zeros(1)))
            return p[:n]
        p = np.zeros(3 * max(n, 1), np.int32)
        c = np.zeros(max(n, 1))
        self.L.tisref_list_get(which, ip(p), fp(c))
        pairs = p[:3 * n].reshape(n, 3)
        return (pairs, c[:n]) if which == 1 else pairs

    def force_term(self, name):
        f = np.zeros((self.sys.nmp, 3))
        _on_big_stack(self.L.tisref_force_term, self.TERMS[name], fp(f))
        return f

    def mt_seed(self, seed):
        """Seed the replica stream the reference's valence pass draws from."""
        self.L.tisref_mt_seed(seed)

    def mt_raw(self, seed, n):
        """The first n 32-bit words of a stream seeded like mt_seed(seed)."""
        r = np.zeros(n, np.int32)
        self.L.tisref_mt_raw(seed, n, ip(r))
        return r.view(np.uint32)

    def hb_state(self):
        nh = max(len(self.sys.hb_mp), 1)
        st = np.zeros(nh, np.int32)
        ene = np.zeros(nh)
        self.L.tisref_hb_state(ip(st), fp(ene))
        return st[:len(self.sys.hb_mp)], ene[:len(self.sys.hb_mp)]

    def energy(self, velo=None):
        e = np.zeros(25)
        if velo is None:
            _on_big_stack(self.L.tisref_energy, fp(np.zeros(3)), 0, fp(e))
        else:
            v = np.ascontiguousarray(velo, dtype=np.float64)
            _on_big_stack(self.L.tisref_energy, fp(v), 1, fp(e))
        return e

    def energy_unit(self, velo=None, style=0):
        nu = self.sys.nunit
        eu = np.zeros(25 * nu * nu)
        if velo is None:
            _on_big_stack(self.L.tisref_energy_unit, fp(np.zeros(3)), 0, style, fp(eu))
        else:
            v = np.ascontiguousarray(velo, dtype=np.float64)
            _on_big_stack(self.L.tisref_energy_unit, fp(v), 1, style, fp(eu))
        return eu.reshape(25, nu, nu).transpose(2, 1, 0)

    def hydro(self, bd, xyz):
        """(D, L) of the reference's simu_hydro_tensors."""
        n = bd.nmp
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        D = np.zeros((3 * n, 3 * n), order="F")
        Lm = np.zeros((3 * n, 3 * n), order="F")
        _on_big_stack(self.L.tisref_hydro, n, fp(x), fp(bd.radius), bd.tempk, bd.viscosity, bd.hydro_tensor,
                      D.ctypes.data_as(_f64p), Lm.ctypes.data_as(_f64p))
        return D, Lm

    def bd_forces(self, bd, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        f = np.zeros_like(x)
        _on_big_stack(self.L.tisref_bd_forces, bd.nmp, fp(x), len(bd.fene_nat), ip(bd.fene_mp), fp(bd.fene_nat),
                      fp(bd.fene_coef), fp(bd.fene_dist2), len(bd.wca_nat2), ip(bd.wca_mp), fp(bd.wca_nat2),
                      fp(bd.wca_coef), bd.cutoff_wca, fp(f))
        return f

    def temperature(self, tempk, ionic):
        hs = np.ascontiguousarray(self.sys.st_hsTm, dtype=np.float64)
        if hs.size == 0:
            hs = np.zeros((1, 3))
        coef, cdist, zp, beta = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        u0 = np.zeros(max(len(self.sys.st_mp), 1))
        self.L.tisref_temperature(tempk, ionic, fp(hs), C.byref(coef), C.byref(cdist),
                                  C.byref(zp), fp(u0), C.byref(beta))
        return coef.value, cdist.value, zp.value, u0[:len(self.sys.st_mp)], beta.value
