"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point that include/tis_hip.h declares, and rejects bad arguments without a
device call."""
import ctypes as C
import os
import re

import pytest

from tis_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("tis_hip.h", "tis_bdhi.h")


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(tis_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("tis_ctx_create", "tis_system_upload", "tis_neighbor", "tis_force_sumup",
                 "tis_energy_sumup", "tis_langevin_init", "tis_langevin_steps", "tis_pairs_get"):
        assert must in names
    for must in ("tis_bdhi_setup", "tis_bdhi_tensors", "tis_bdhi_steps", "tis_bdhi_factor_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_rejected():
    lib = engine.load_library()
    assert lib.tis_ctx_destroy(None) < 0
    assert lib.tis_neighbor(None, 0) < 0
    assert lib.tis_force_sumup(None, 0, None) < 0
    assert lib.tis_langevin_steps(None, 1, 0) < 0
    assert lib.tis_bdhi_tensors(None) < 0
    assert lib.tis_bdhi_steps(None, C.c_double(0.05), 1, 1, 0) < 0
    assert lib.tis_system_origin(None, 0, 0, 0) < 0
    assert lib.tis_energy_sumup_set(None, 1, None, None) < 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(engine, "LIBPATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(engine, "_LIB", None)
    with pytest.raises(engine.TisError):
        engine.load_library()


def test_fortran_module_binds_every_entry_point():
    """module tis_hip (tis_amd/fortran/tis_hip.F90) has a bind(C) interface
    for every function the headers declare."""
    src = open(os.path.join(ROOT, "tis_amd", "fortran", "tis_hip.F90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(tis_[a-z0-9_]+)"\)', src))
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_validate_rejects_out_of_range_indices():
    """tis_system_validate (run first by tis_system_upload) names the first
    index outside its bounds; a 1-based bond index, a site past nhbsite or a
    tertiary stack mapped past nst is refused without a device call."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from systems import hairpin64, tstack_hairpin
    s, _ = hairpin64()
    engine.validate([s])
    bad = s.bd_mp.copy()
    bad[-1, 1] = s.nmp  # a 1-based index slip
    s.bd_mp = bad
    with pytest.raises(engine.TisError, match="bd_mp"):
        engine.validate([s])
    s, _ = hairpin64()
    site = s.hb_site.copy()
    site[0, 0] = len(s.hbsite_valence)
    s.hb_site = site
    with pytest.raises(engine.TisError, match="hb_site"):
        engine.validate([s])
    t, _ = tstack_hairpin()
    engine.validate([t])
    m = t.tst2st.copy()
    m[0, 0] = len(t.st_mp)
    t.tst2st = m
    with pytest.raises(engine.TisError, match="tst2st"):
        engine.validate([hairpin64()[0], t])
