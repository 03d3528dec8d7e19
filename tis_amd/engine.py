"""Host-side mirror of the reference's per-replica call sites over the HIP
C-ABI (include/tis_hip.h, libtis_hip.so).

``Engine`` holds one context with a batch of systems resident in HBM and
exposes the reference's entry points by name:

  neighbor(isys)            neighbor(irep)                  src/neighbor.F90:13
  force_sumup(isys)         force_sumup(force_mp, irep)     src/force_sumup.F90:15
  energy_sumup(isys)        energy_sumup(irep, velo, E, ..) src/energy_sumup.F90:17
  langevin_init / _steps    time_integral (Langevin branch) src/time_integral.F90:149-189

The native library must be present: there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .sysdesc import CSystem, System, E_MAX

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtis_hip.so")

MODE_GRAPH, MODE_EAGER, MODE_PLAIN = 0, 1, 2
LIST_EXV, LIST_ELE, LIST_HB = 0, 1, 2
CNT_NAMES = ("clamp_exv", "clamp_ele", "hb_deleted", "rebuilds", "steps", "nb_overflow", "run_completed",
             "resumes")

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_LIB = None


class TisError(RuntimeError):
    pass


def load_library():
    """Load libtis_hip.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise TisError("libtis_hip.so not built: run __graft_entry__.build() (no CPU fallback)")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    L.tis_ctx_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.tis_ctx_destroy.argtypes = [vp]
    L.tis_last_error.restype = C.c_char_p
    L.tis_last_error.argtypes = [vp]
    L.tis_system_upload.argtypes = [vp, C.c_int32, C.POINTER(CSystem), C.POINTER(C.c_uint32), C.c_uint64]
    L.tis_system_validate.argtypes = [C.c_int32, C.POINTER(CSystem), C.c_char_p, C.c_int64]
    L.tis_set_system_T.argtypes = [vp, C.c_int32, C.POINTER(CSystem)]
    L.tis_list_capacity.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.tis_term_status.argtypes = [vp, C.c_int32, _i32p, _i32p]
    L.tis_transfer_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.tis_system_append.argtypes = [vp, C.c_int32, C.POINTER(CSystem), C.POINTER(C.c_uint32), _i32p]
    L.tis_system_retire.argtypes = [vp, C.c_int32, _i32p]
    L.tis_arena_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    L.tis_tune.argtypes = [vp, C.c_int32, _i32p, _f64p]
    L.tis_launch_shape.argtypes = [vp, _i32p, _i32p]
    L.tis_state_set.argtypes = [vp, C.c_int32, C.c_int32, _f64p]
    L.tis_state_get.argtypes = [vp, C.c_int32, C.c_int32, _f64p]
    L.tis_neighbor.argtypes = [vp, C.c_int32]
    L.tis_pairs_get.argtypes = [vp, C.c_int32, C.c_int32, _i32p, _f64p, C.c_int64, C.POINTER(C.c_int64)]
    L.tis_force_sumup.argtypes = [vp, C.c_int32, _f64p]
    L.tis_force_get.argtypes = [vp, C.c_int32, _f64p]
    L.tis_energy_sumup.argtypes = [vp, C.c_int32, _f64p]
    L.tis_energy_sumup_set.argtypes = [vp, C.c_int32, _i32p, _f64p]
    L.tis_energy_unit.argtypes = [vp, C.c_int32, C.c_int32, _f64p]
    L.tis_langevin_init.argtypes = [vp, C.c_double, C.c_int32, C.c_uint64, C.c_int32]
    L.tis_langevin_steps.argtypes = [vp, C.c_int64, C.c_int32]
    L.tis_system_origin.argtypes = [vp, C.c_int32, C.c_uint64, C.c_int32]
    L.tis_device_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.tis_label_define.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(CSystem)]
    L.tis_label_apply.argtypes = [vp, C.c_int32, _i32p, _i32p, _f64p]
    L.tis_status_counters.argtypes = [vp, C.POINTER(C.c_int64)]
    L.tis_synchronize.argtypes = [vp]
    L.tis_last_timing.argtypes = [vp, _f64p, _f64p, _f64p, C.POINTER(C.c_int64)]
    L.tis_last_force_ms.argtypes = [vp, _f64p]
    for fn in ("tis_ctx_create", "tis_ctx_destroy", "tis_system_upload", "tis_system_validate", "tis_set_system_T",
               "tis_list_capacity", "tis_term_status", "tis_transfer_bytes", "tis_system_append",
               "tis_system_retire", "tis_arena_stats", "tis_tune", "tis_launch_shape",
               "tis_state_set",
               "tis_state_get", "tis_neighbor", "tis_pairs_get", "tis_force_sumup", "tis_force_get",
               "tis_energy_sumup", "tis_energy_sumup_set", "tis_energy_unit", "tis_langevin_init", "tis_langevin_steps", "tis_status_counters",
               "tis_system_origin", "tis_device_bytes", "tis_label_define", "tis_label_apply",
               "tis_synchronize", "tis_last_timing", "tis_last_force_ms"):
        getattr(L, fn).restype = C.c_int32
    _LIB = L
    return L


def _fp(a):
    return a.ctypes.data_as(_f64p)


def validate(systems) -> None:
    """tis_system_validate: raise TisError naming the first index out of
    range in a batch description (no device call)."""
    L = load_library()
    cs = (CSystem * len(systems))(*[s.cstruct() for s in systems])
    msg = C.create_string_buffer(512)
    rc = L.tis_system_validate(len(systems), cs, msg, len(msg))
    if rc != 0:
        raise TisError(f"invalid system ({rc}): {msg.value.decode()}")


class Engine:
    """A batch of DT15 systems on one GPU."""

    def __init__(self, systems: list[System], keys=None, seed: int = 1, device: int = 0):
        self.L = load_library()
        self.ctx = C.c_void_p()
        self._chk(self.L.tis_ctx_create(device, C.byref(self.ctx)), "tis_ctx_create")
        self.upload(systems, keys, seed)

    def upload(self, systems: list[System], keys=None, seed: int = 1):
        """Replace the batch (the context, its stream and library state stay)."""
        self.systems = list(systems)
        self._cs = (CSystem * len(self.systems))(*[s.cstruct() for s in self.systems])
        k = None
        if keys is not None:
            if len(keys) != len(self.systems):
                raise ValueError("one key per system")
            self._keys = (C.c_uint32 * len(keys))(*[int(v) for v in keys])
            k = self._keys
        self._chk(self.L.tis_system_upload(self.ctx, len(self.systems), self._cs, k, seed), "upload")
        self.live = set(range(len(self.systems)))

    def append(self, systems: list[System], keys=None) -> list[int]:
        """Add systems to the batch (tis_system_append): only their own
        topology is uploaded; returns their indices."""
        n = len(systems)
        cs = (CSystem * n)(*[s.cstruct() for s in systems])
        k = None if keys is None else (C.c_uint32 * n)(*[int(v) for v in keys])
        out = (C.c_int32 * n)()
        self._chk(self.L.tis_system_append(self.ctx, n, cs, k, out), "tis_system_append")
        self.systems.extend(systems)
        self.live.update(int(v) for v in out)
        return [int(v) for v in out]

    def retire(self, isys):
        """Drop systems from every launch (tis_system_retire)."""
        a = np.ascontiguousarray(isys, dtype=np.int32)
        self._chk(self.L.tis_system_retire(self.ctx, len(a), a.ctypes.data_as(_i32p)), "tis_system_retire")
        self.live.difference_update(int(v) for v in a)

    def tune(self, reps: int = 20) -> dict:
        """tis_tune: time the force pass under each launch shape, keep the
        fastest; returns the choice and the per-candidate times."""
        shape = (C.c_int32 * 3)()
        ms = (C.c_double * 8)()
        self._chk(self.L.tis_tune(self.ctx, int(reps), shape, ms), "tis_tune")
        names = [f"ppw{p}_tpl{t}_{'offdiag_first' if o else 'system_order'}"
                 for o in (1, 0) for t in (1, 2) for p in (1, 2)]
        return {"pairs_per_wave": shape[0], "terms_per_lane": shape[1], "offdiag_first": bool(shape[2]),
                "candidates_ms": dict(zip(names, [round(v, 5) for v in ms]))}

    def launch_shape(self, shape=None):
        """Set (ppw, tpl, offdiag_first) when given; return the current one."""
        get = (C.c_int32 * 3)()
        st = None if shape is None else (C.c_int32 * 3)(*[int(v) for v in shape])
        self._chk(self.L.tis_launch_shape(self.ctx, st, get), "tis_launch_shape")
        return tuple(get)

    def arena_stats(self) -> dict:
        a = (C.c_int64 * 6)()
        self._chk(self.L.tis_arena_stats(self.ctx, a), "tis_arena_stats")
        return dict(zip(("live_systems", "live_beads", "stored_beads", "bead_capacity", "layouts", "indices"),
                        list(a)))

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.tis_last_error(self.ctx).decode() if self.ctx else ""
            raise TisError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.ctx:
            self.L.tis_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def set_state(self, isys, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (self.systems[isys].nmp, 3):
            raise ValueError("state must be (N,3)")
        self._chk(self.L.tis_state_set(self.ctx, isys, which, _fp(a)), "tis_state_set")

    def get_state(self, isys, which):
        a = np.zeros((self.systems[isys].nmp, 3))
        self._chk(self.L.tis_state_get(self.ctx, isys, which, _fp(a)), "tis_state_get")
        return a

    def set_coords(self, isys, xyz):
        self.set_state(isys, 0, xyz)

    def coords(self, isys):
        return self.get_state(isys, 0)

    def set_temperature_system(self, isys, system: System):
        cs = system.cstruct()
        self._chk(self.L.tis_set_system_T(self.ctx, isys, C.byref(cs)), "tis_set_system_T")
        self.systems[isys] = system

    # ---- reference entry points
    def neighbor(self, isys=-1):
        self._chk(self.L.tis_neighbor(self.ctx, isys), "tis_neighbor")

    def pairs(self, isys, which):
        n = C.c_int64()
        self.L.tis_pairs_get(self.ctx, isys, which, None, None, 0, C.byref(n))
        cnt = n.value
        if which == LIST_HB:
            buf = np.zeros(max(cnt, 1), np.int32)
            self._chk(self.L.tis_pairs_get(self.ctx, isys, which, buf.ctypes.data_as(_i32p), None, cnt,
                                           C.byref(n)), "tis_pairs_get")
            return buf[:cnt]
        buf = np.zeros(3 * max(cnt, 1), np.int32)
        coef = np.zeros(max(cnt, 1))
        self._chk(self.L.tis_pairs_get(self.ctx, isys, which, buf.ctypes.data_as(_i32p), _fp(coef), cnt,
                                       C.byref(n)), "tis_pairs_get")
        pairs = buf[:3 * cnt].reshape(cnt, 3)
        return (pairs, coef[:cnt]) if which == LIST_ELE else pairs

    def force_sumup(self, isys):
        f = np.zeros((self.systems[isys].nmp, 3))
        self._chk(self.L.tis_force_sumup(self.ctx, isys, _fp(f)), "tis_force_sumup")
        return f

    def force_all(self):
        """Forces of every live system (None for retired ones)."""
        self._chk(self.L.tis_force_sumup(self.ctx, -1, None), "tis_force_sumup")
        return [self.get_state(i, 3) if i in self.live else None for i in range(len(self.systems))]

    def energy_sumup(self, isys):
        e = np.zeros(E_MAX)
        self._chk(self.L.tis_energy_sumup(self.ctx, isys, _fp(e)), "tis_energy_sumup")
        return e

    def energy_unit(self, isys, style=0):
        """energy_unit (nunit, nunit, E_MAX) of system isys, indexed [iunit, junit, type]."""
        nu = self.systems[isys].nunit
        eu = np.zeros(E_MAX * nu * nu)
        self._chk(self.L.tis_energy_unit(self.ctx, isys, style, _fp(eu)), "tis_energy_unit")
        return eu.reshape(E_MAX, nu, nu).transpose(2, 1, 0)

    def energy_all(self):
        """(nsys, E_MAX) energies of every system in one launch."""
        e = np.zeros((len(self.systems), E_MAX))
        self._chk(self.L.tis_energy_sumup(self.ctx, -1, _fp(e)), "tis_energy_sumup")
        return e

    def energy_systems(self, isys):
        """(len(isys), E_MAX) energies of the systems isys, in that order, in
        one launch."""
        a = np.ascontiguousarray(isys, dtype=np.int32)
        e = np.zeros((len(a), E_MAX))
        self._chk(self.L.tis_energy_sumup_set(self.ctx, len(a), a.ctypes.data_as(_i32p), _fp(e)),
                  "tis_energy_sumup_set")
        return e

    # ---- Langevin
    def system_origin(self, isys, step_origin, init_accel):
        self._chk(self.L.tis_system_origin(self.ctx, isys, int(step_origin), 1 if init_accel else 0),
                  "tis_system_origin")

    def langevin_init(self, tstep=0.05, rebuild_every=10, step0=0, init_accel=True):
        """init_accel: True, False, or None for the per-system flags of system_origin."""
        flag = -1 if init_accel is None else (1 if init_accel else 0)
        self._chk(self.L.tis_langevin_init(self.ctx, tstep, rebuild_every, step0, flag), "tis_langevin_init")

    def langevin_steps(self, nsteps, mode=MODE_GRAPH):
        self._chk(self.L.tis_langevin_steps(self.ctx, nsteps, mode), "tis_langevin_steps")

    def synchronize(self):
        self._chk(self.L.tis_synchronize(self.ctx), "tis_synchronize")

    def label_define(self, label, isys_template, system: System):
        cs = system.cstruct()
        self._chk(self.L.tis_label_define(self.ctx, label, isys_template, C.byref(cs)), "tis_label_define")
        if not hasattr(self, "_labels"):
            self._labels = {}
        self._labels[label] = system

    def label_apply(self, isys, labels, vscale=None):
        """Systems isys[k] take label labels[k]; velocities times vscale[k]."""
        n = len(isys)
        a = np.ascontiguousarray(isys, dtype=np.int32)
        b = np.ascontiguousarray(labels, dtype=np.int32)
        v = None if vscale is None else np.ascontiguousarray(vscale, dtype=np.float64)
        self._chk(self.L.tis_label_apply(self.ctx, n, a.ctypes.data_as(_i32p), b.ctypes.data_as(_i32p),
                                         None if v is None else _fp(v)), "tis_label_apply")
        for k, lab in zip(a, b):
            self.systems[int(k)] = self._labels[int(lab)]

    def device_bytes(self):
        n = C.c_int64()
        self._chk(self.L.tis_device_bytes(self.ctx, C.byref(n)), "tis_device_bytes")
        return n.value

    def counters(self):
        a = (C.c_int64 * 8)()
        self._chk(self.L.tis_status_counters(self.ctx, a), "tis_status_counters")
        return dict(zip(CNT_NAMES, list(a)[:len(CNT_NAMES)]))

    def transfer_bytes(self) -> dict:
        """Payload bytes since creation (tis_transfer_bytes): topology and
        tables uploaded, state uploaded, state downloaded, state carried on
        the device by a new layout."""
        a = (C.c_int64 * 4)()
        self._chk(self.L.tis_transfer_bytes(self.ctx, a), "tis_transfer_bytes")
        return dict(zip(("h2d_topology", "h2d_state", "d2h_state", "d2d_state"), list(a)))

    def term_status(self, isys):
        """(hb_status[nhb], st_status[nst]) as bool arrays (tis_term_status)."""
        s = self.systems[isys]
        hb = np.zeros(max(len(s.hb_mp), 1), np.int32)
        st = np.zeros(max(len(s.st_mp), 1), np.int32)
        self._chk(self.L.tis_term_status(self.ctx, isys, hb.ctypes.data_as(_i32p), st.ctypes.data_as(_i32p)),
                  "tis_term_status")
        return hb[:len(s.hb_mp)] != 0, st[:len(s.st_mp)] != 0

    def list_capacity(self, cap: int = 0) -> int:
        """Entries per bead of the exv lists; cap > 0 sets it (tis_list_capacity)."""
        out = C.c_int32()
        self._chk(self.L.tis_list_capacity(self.ctx, int(cap), C.byref(out)), "tis_list_capacity")
        return out.value

    def last_timing(self):
        t, m, p = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int64()
        self._chk(self.L.tis_last_timing(self.ctx, C.byref(t), C.byref(m), C.byref(p), C.byref(n)),
                  "tis_last_timing")
        e = C.c_double()
        self._chk(self.L.tis_last_force_ms(self.ctx, C.byref(e)), "tis_last_force_ms")
        return {"total_ms": t.value, "main_ms": m.value, "pre_ms": p.value, "force_ms": e.value, "nlaunch": n.value}


class BDEngine:
    """Brownian dynamics with hydrodynamic interaction for one bead system
    (include/tis_bdhi.h) on its own context."""

    def __init__(self, system, key: int = 0, seed: int = 1, device: int = 0):
        self.L = load_library()
        L = self.L
        vp = C.c_void_p
        L.tis_bdhi_setup.argtypes = [vp, C.c_void_p, C.c_uint32, C.c_uint64]
        L.tis_bdhi_state_set.argtypes = [vp, C.c_int32, _f64p]
        L.tis_bdhi_state_get.argtypes = [vp, C.c_int32, _f64p]
        L.tis_bdhi_forces.argtypes = [vp, _f64p]
        L.tis_bdhi_tensors.argtypes = [vp]
        L.tis_bdhi_tensor_get.argtypes = [vp, C.c_int32, _f64p]
        L.tis_bdhi_factor_error.argtypes = [vp, _f64p]
        L.tis_bdhi_steps.argtypes = [vp, C.c_double, C.c_int64, C.c_int32, C.c_uint64]
        L.tis_bdhi_timing.argtypes = [vp, _f64p, _f64p, _f64p]
        for fn in ("tis_bdhi_setup", "tis_bdhi_state_set", "tis_bdhi_state_get", "tis_bdhi_forces",
                   "tis_bdhi_tensors", "tis_bdhi_tensor_get", "tis_bdhi_factor_error", "tis_bdhi_steps",
                   "tis_bdhi_timing"):
            getattr(L, fn).restype = C.c_int32
        self.ctx = C.c_void_p()
        self._chk(L.tis_ctx_create(device, C.byref(self.ctx)), "tis_ctx_create")
        self.system = system
        self._cs = system.cstruct()
        self._chk(L.tis_bdhi_setup(self.ctx, C.cast(C.pointer(self._cs), C.c_void_p), key, seed), "tis_bdhi_setup")

    _chk = Engine._chk
    close = Engine.close
    __del__ = Engine.__del__

    @property
    def n3(self):
        return 3 * self.system.nmp

    def set_coords(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        if a.shape != (self.system.nmp, 3):
            raise ValueError("coordinates must be (N,3)")
        self._chk(self.L.tis_bdhi_state_set(self.ctx, 0, _fp(a)), "tis_bdhi_state_set")

    def coords(self):
        a = np.zeros((self.system.nmp, 3))
        self._chk(self.L.tis_bdhi_state_get(self.ctx, 0, _fp(a)), "tis_bdhi_state_get")
        return a

    def forces(self):
        a = np.zeros((self.system.nmp, 3))
        self._chk(self.L.tis_bdhi_forces(self.ctx, _fp(a)), "tis_bdhi_forces")
        return a

    def tensors(self):
        self._chk(self.L.tis_bdhi_tensors(self.ctx), "tis_bdhi_tensors")

    def tensor(self, which):
        """0: D, 1: L, as (3n, 3n) arrays (column-major on the device)."""
        a = np.zeros((self.n3, self.n3), order="F")
        self._chk(self.L.tis_bdhi_tensor_get(self.ctx, which, a.ctypes.data_as(_f64p)), "tis_bdhi_tensor_get")
        return a

    def factor_error(self):
        e = C.c_double()
        self._chk(self.L.tis_bdhi_factor_error(self.ctx, C.byref(e)), "tis_bdhi_factor_error")
        return e.value

    def steps(self, nsteps, tstep=0.05, refactor_every=10, step0=0):
        self._chk(self.L.tis_bdhi_steps(self.ctx, tstep, nsteps, refactor_every, step0), "tis_bdhi_steps")

    def timing(self):
        t, f, k = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.tis_bdhi_timing(self.ctx, C.byref(t), C.byref(f), C.byref(k)), "tis_bdhi_timing")
        return {"total_ms": t.value, "fill_ms": f.value, "factor_ms": k.value}
